// gfx950 kernels of the InsPIRe respond path.
//
//   k_expand_mask    w_mask rows from the CRS seed (expand_mask, inspiring2.rs:866-887),
//                    transformed and scaled: w' = n^-1 NTT(w_mask_k)
//   k_seat           y' = n^-1 NTT(y_body_k) (PackingKeySeat::build, session.rs:600-624)
//   k_transpose      records -> column-major scratch for the store build
//   k_store_build    S'_i = n^-1 S_i (ShardPackingStore::build + rotation_sums,
//                    shard_store.rs:86-156, inspiring2.rs:964-1004)
//   k_column_tags    keyed column tags (column_tag, shard_store.rs:393-410)
//   k_scan           DB scan: b fold + tag recompute (check_columns_and_fold_b /
//                    fold_stripe, shard_store.rs:197-265, 414-427)
//   k_chain          per query: seed expansion, step-1 gather-multiply, the backward
//                    recursion with the online gather fused in, the final transforms
//                    and the optional 36-bit modulus switch (CheckedColumns::packing_offline
//                    shard_store.rs:282-342, fold_rotations_backward inspiring2.rs:1014-1085,
//                    packing_online_gather :1393-1485, mod_switch_poly mod_switch.rs:148-175)
#include <hip/hip_runtime.h>

#include "chacha.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "modq.hpp"
#include "ntt.hpp"

namespace inspire {

// ---------------------------------------------------------------------------
template <int LOGN>
__global__ __launch_bounds__(kThreads) void k_expand_mask(const uint8_t *seed, uint64_t stream, int dropped,
                                                           const uint64_t *psi, uint64_t inv_n, uint64_t *out) {
    constexpr int N = 1 << LOGN;
    __shared__ uint64_t x[LdsSize<LOGN>::value];
    const int row = blockIdx.x + dropped;  // expansion row; rows below `dropped` are drawn and skipped
    uint32_t key[8];
    seed_key(seed, key);
    for (int blk = threadIdx.x; blk < N / 8; blk += kThreads) {
        uint32_t w[16];
        chacha20_block(key, (uint64_t)row * (N / 8) + blk, stream, w);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            uint64_t v = (uint64_t)w[2 * u] | (uint64_t)w[2 * u + 1] << 32;
            x[lds_ix(blk * 8 + u)] = v % kQ;
        }
    }
    __syncthreads();
    ntt_forward<LOGN>(x, psi);
    for (int j = threadIdx.x; j < N; j += kThreads) out[(size_t)blockIdx.x * N + j] = mulmod(x[lds_ix(j)], inv_n);
}

template <int LOGN>
__global__ __launch_bounds__(kThreads) void k_seat(const uint64_t *y, const uint64_t *psi, uint64_t inv_n,
                                                    uint64_t *out) {
    constexpr int N = 1 << LOGN;
    __shared__ uint64_t x[LdsSize<LOGN>::value];
    for (int j = threadIdx.x; j < N; j += kThreads) x[lds_ix(j)] = y[(size_t)blockIdx.x * N + j];
    __syncthreads();
    ntt_forward<LOGN>(x, psi);
    for (int j = threadIdx.x; j < N; j += kThreads) out[(size_t)blockIdx.x * N + j] = mulmod(x[lds_ix(j)], inv_n);
}

__global__ void k_transpose(const uint16_t *rec, int n, int gamma, uint16_t *col) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * gamma) return;
    int m = (int)(idx / gamma), j = (int)(idx % gamma);
    col[(size_t)j * n + m] = rec[idx];
}

// S'_i[k] = (n*gamma)^-1 * NTT(P_i)[tab_{g^i}[k]], where P_i = sum_j X^(j*g^(n-i)) h_j is
// formed in the coefficient domain (each term is a signed negacyclic shift of a column).
template <int LOGN>
__global__ __launch_bounds__(kThreads) void k_store_build(CtxDev c, const uint16_t *col, uint64_t *store) {
    constexpr int N = 1 << LOGN;
    __shared__ uint64_t x[LdsSize<LOGN>::value];
    const int i = blockIdx.x;
    const uint32_t back = c.gpow_back[i];
    const uint32_t two_n = 2u * N;
    for (int k = threadIdx.x; k < N; k += kThreads) {
        int32_t sum = 0;
        for (int j = 0; j < c.gamma; j++) {
            uint32_t s = (uint32_t)(((uint64_t)j * back) % two_n);
            uint32_t r = s & (N - 1);
            bool neg = s >= (uint32_t)N;
            int src = (k - (int)r) & (N - 1);
            bool wrap = k < (int)r;
            int32_t h = col[(size_t)j * N + src];
            sum += (neg != wrap) ? -h : h;
        }
        x[lds_ix(k)] = sum >= 0 ? (uint64_t)sum : kQ - (uint64_t)(-(int64_t)sum);
    }
    __syncthreads();
    ntt_forward<LOGN>(x, c.psi);
    const uint32_t t = c.gpow[i];
    for (int k = threadIdx.x; k < N; k += kThreads)
        store[(size_t)i * N + k] = mulmod(x[lds_ix(automorph_slot(t, k, LOGN))], c.inv_n_gamma);
}

__global__ __launch_bounds__(kThreads) void k_column_tags(const uint16_t *col, const uint64_t *key, int n,
                                                           uint64_t *tags) {
    __shared__ uint64_t red_lo[kThreads], red_hi[kThreads];
    const int j = blockIdx.x;
    uint64_t lo = 0, hi = 0;
    for (int k = threadIdx.x; k < n; k += kThreads) {
        uint64_t h = col[(size_t)j * n + k];
        lo += h * (key[k] & 0xffffffffu);
        hi += h * (key[k] >> 32);
    }
    red_lo[threadIdx.x] = lo;
    red_hi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red_lo[threadIdx.x] += red_lo[threadIdx.x + s];
            red_hi[threadIdx.x] += red_hi[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint64_t L = red_lo[0], H = red_hi[0];  // value = H*2^32 + L
        uint64_t vlo = L + (H << 32);
        uint64_t vhi = (H >> 32) + (vlo < L);
        tags[j] = mod_p61(vhi, vlo);
    }
}

// DB scan. Block (query, split) covers rows [split*n/ks, (split+1)*n/ks) of the
// query's shard and every column; per column it accumulates sum_k brev[k]*h[k] and
// sum_k key[k]*h[k] as two 32-bit halves in u64 (no overflow for n <= 4096).
__global__ __launch_bounds__(kThreads) void k_scan(CtxDev c, BatchDev bt) {
    const int qi = blockIdx.x, split = blockIdx.y;
    const int n = c.n, gamma = c.gamma;
    const ShardDev sh = c.shards[bt.shard_ids[qi]];
    const uint64_t *b = bt.b + (size_t)qi * n;
    const int rows = n / bt.ks, r0 = split * rows;
    const int cp = gamma < kThreads ? gamma : kThreads;
    const int lanes = kThreads / cp;
    __shared__ uint64_t red[kThreads][4];
    for (int jb = 0; jb < gamma; jb += cp) {
        const int j = jb + (int)threadIdx.x % cp, lane = (int)threadIdx.x / cp;
        uint64_t bl = 0, bh = 0, tl = 0, th = 0;
        if (lane < lanes && j < gamma) {
            for (int k = r0 + lane; k < r0 + rows; k += lanes) {
                uint64_t h = sh.records[(size_t)k * gamma + j];
                uint64_t bk = k == 0 ? b[0] : (b[n - k] ? kQ - b[n - k] : 0);
                uint64_t key = sh.tag_key[k];
                bl += h * (bk & 0xffffffffu);
                bh += h * (bk >> 32);
                tl += h * (key & 0xffffffffu);
                th += h * (key >> 32);
            }
        }
        red[threadIdx.x][0] = bl;
        red[threadIdx.x][1] = bh;
        red[threadIdx.x][2] = tl;
        red[threadIdx.x][3] = th;
        __syncthreads();
        if ((int)threadIdx.x < cp && jb + (int)threadIdx.x < gamma) {
            uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int l = 0; l < lanes; l++) {
                int t = l * cp + threadIdx.x;
                s0 += red[t][0]; s1 += red[t][1]; s2 += red[t][2]; s3 += red[t][3];
            }
            uint64_t *o = bt.scan_part + (((size_t)qi * bt.ks + split) * gamma + jb + threadIdx.x) * 4;
            o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// One workgroup answers one query.
template <int LOGN, int ROWS, bool EXACT>
__global__ __launch_bounds__(kThreads) void k_chain(CtxDev c, BatchDev bt) {
    constexpr int N = 1 << LOGN;
    constexpr int SLOTS = N >= kThreads ? N / kThreads : 1;
    __shared__ uint64_t X[LdsSize<LOGN>::value];   // n^-1 R_{i+1}, NTT form
    __shared__ uint64_t C[LdsSize<LOGN>::value];   // digit 0 / scratch
    __shared__ uint64_t D[LdsSize<LOGN>::value];   // digit 1
    __shared__ uint64_t A[LdsSize<LOGN>::value];   // NTT(a)
    __shared__ uint64_t bpoly[1024];
    __shared__ int bad_column;
    __shared__ int rejected;

    const int qi = blockIdx.x, tid = threadIdx.x, gamma = c.gamma;
    const ShardDev sh = c.shards[bt.shard_ids[qi]];

    // --- finish the scan: b_poly and the tag check, before any packing work
    if (tid == 0) { bad_column = 0x7fffffff; rejected = 0; }
    __syncthreads();
    for (int j = tid; j < gamma; j += kThreads) {
        uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int s = 0; s < bt.ks; s++) {
            const uint64_t *p = bt.scan_part + (((size_t)qi * bt.ks + s) * gamma + j) * 4;
            s0 += p[0]; s1 += p[1]; s2 += p[2]; s3 += p[3];
        }
        uint64_t lo = s0 + (s1 << 32), hi = (s1 >> 32) + (lo < s0);
        bpoly[j] = reduce120(hi, lo);
        uint64_t tlo = s2 + (s3 << 32), thi = (s3 >> 32) + (tlo < s2);
        if (mod_p61(thi, tlo) != sh.tags[j]) atomicMin(&bad_column, j);
    }
    __syncthreads();
    if (bad_column != 0x7fffffff) {
        if (tid == 0) bt.status[qi] = bad_column + 1;
        return;
    }

    // --- a = gen_range(0..q) draws from ChaCha20(seed); one keystream block per lane
    if (bt.a_in) {
        for (int j = tid; j < N; j += kThreads) A[lds_ix(j)] = bt.a_in[(size_t)qi * N + j];
        __syncthreads();
    } else {
        uint32_t key[8];
        seed_key(bt.seeds + (size_t)qi * 32, key);
        for (int blk = tid; blk < N / 8; blk += kThreads) {
            uint32_t w[16];
            chacha20_block(key, (uint64_t)blk, 0, w);
#pragma unroll
            for (int u = 0; u < 8; u++) {
                uint64_t v = (uint64_t)w[2 * u] | (uint64_t)w[2 * u + 1] << 32, a;
                if (!uniform_q_draw(v, a)) rejected = 1;
                A[lds_ix(blk * 8 + u)] = a;
            }
        }
        __syncthreads();
        if (rejected) {
            // A draw fell outside the acceptance zone (probability ~2^-46 per draw):
            // replay the stream serially with rejection, as the reference's RNG does.
            if (tid == 0) {
                uint64_t blk = 0;
                int word = 16, got = 0;
                uint32_t w[16];
                while (got < N) {
                    uint32_t lo32, hi32;
                    if (word >= 16) { chacha20_block(key, blk++, 0, w); word = 0; }
                    lo32 = w[word++];
                    if (word >= 16) { chacha20_block(key, blk++, 0, w); word = 0; }
                    hi32 = w[word++];
                    uint64_t a;
                    if (uniform_q_draw((uint64_t)lo32 | (uint64_t)hi32 << 32, a)) A[lds_ix(got++)] = a;
                }
            }
            __syncthreads();
        }
    }
    ntt_forward<LOGN>(A, c.psi);

    // --- step 1 for the last rotation: X = NTT(a)[tab_{gamma-1}] * S'_{gamma-1}
    uint64_t acc[SLOTS];
#pragma unroll
    for (int v = 0; v < SLOTS; v++) {
        acc[v] = 0;
        const int j = tid + v * kThreads;
        if (j < N) {
            uint32_t src = automorph_slot(c.gpow[gamma - 1], j, LOGN);
            X[lds_ix(j)] = mulmod(A[lds_ix(src)], sh.store[(size_t)(gamma - 1) * N + j]);
        }
    }
    __syncthreads();

    // --- backward recursion with the online gather
    for (int i = gamma - 2; i >= 0; i--) {
        for (int j = tid; j < N; j += kThreads) C[lds_ix(j)] = X[lds_ix(j)];
        __syncthreads();
        ntt_inverse_noscale<LOGN>(C, c.psi_inv);  // coefficients of R_{i+1}, canonical
        for (int j = tid; j < N; j += kThreads) {
            uint64_t v = C[lds_ix(j)];
            if (EXACT) {
                C[lds_ix(j)] = v & 0xfffff;
                D[lds_ix(j)] = (v >> 20) & 0xfffff;
            } else {
                uint64_t r = (v + (1ULL << 19)) >> 20;  // rounded digits, dropped = 1
                C[lds_ix(j)] = r & 0xfffff;
                D[lds_ix(j)] = r >> 20;
            }
        }
        __syncthreads();
        ntt_forward<LOGN>(C, c.psi);
        ntt_forward<LOGN>(D, c.psi);
        const uint32_t t = c.gpow[i];
        const uint64_t *S = sh.store + (size_t)i * N;
#pragma unroll
        for (int v = 0; v < SLOTS; v++) {
            const int j = tid + v * kThreads;
            if (j >= N) break;
            const uint32_t src = automorph_slot(t, j, LOGN);
            uint64_t d[ROWS];
            d[0] = C[lds_ix(j)];
            d[1] = D[lds_ix(j)];
            if constexpr (ROWS == 3) {
                Acc128 l;
                l.mac(c.kx, X[lds_ix(j)]);
                l.mac(c.k0, d[0]);
                l.mac(c.k1, d[1]);
                d[2] = l.reduce();
            }
            Acc128 r, o;
            r.mac(A[lds_ix(src)], S[j]);
            o.add(acc[v]);
#pragma unroll
            for (int k = 0; k < ROWS; k++) {
                r.mac(d[k], c.wprime[(size_t)k * N + src]);
                o.mac(d[k], bt.yprime[(size_t)k * N + src]);
            }
            acc[v] = o.reduce();
            X[lds_ix(j)] = r.reduce();
        }
        __syncthreads();
    }

    // --- a_hat = iNTT(R_0); b = b_poly + iNTT(sum); optional modulus switch
    for (int j = tid; j < N; j += kThreads) C[lds_ix(j)] = X[lds_ix(j)];
#pragma unroll
    for (int v = 0; v < SLOTS; v++) {
        const int j = tid + v * kThreads;
        if (j < N) D[lds_ix(j)] = acc[v];
    }
    __syncthreads();
    ntt_inverse_noscale<LOGN>(C, c.psi_inv);
    ntt_inverse_noscale<LOGN>(D, c.psi_inv);
    const uint64_t qn = bt.switch_modulus;
    for (int j = tid; j < N; j += kThreads) {
        uint64_t a = C[lds_ix(j)];
        bt.a_out[(size_t)qi * N + j] = qn ? mod_switch_coeff(a, qn) : a;
    }
    for (int j = tid; j < gamma; j += kThreads) {
        uint64_t bv = addmod(bpoly[j], D[lds_ix(j)]);
        bt.b_out[(size_t)qi * gamma + j] = qn ? mod_switch_coeff(bv, qn) : bv;
    }
    if (tid == 0) bt.status[qi] = 0;
}

// ---------------------------------------------------------------------------
// launchers

template <int LOGN>
static hipError_t launch_chain(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    if (c.rows == 3)
        hipLaunchKernelGGL((k_chain<LOGN, 3, true>), dim3(b.count), dim3(kThreads), 0, s, c, b);
    else
        hipLaunchKernelGGL((k_chain<LOGN, 2, false>), dim3(b.count), dim3(kThreads), 0, s, c, b);
    return hipGetLastError();
}

hipError_t launch_respond(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    if (b.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan, dim3(b.count, b.ks), dim3(kThreads), 0, s, c, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (c.logn) {
        case 8: return launch_chain<8>(c, b, s);
        case 11: return launch_chain<11>(c, b, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_scan_only(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(b.count, b.ks), dim3(kThreads), 0, s, c, b);
    return hipGetLastError();
}

hipError_t launch_chain_only(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    switch (c.logn) {
        case 8: return launch_chain<8>(c, b, s);
        case 11: return launch_chain<11>(c, b, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_expand_mask(int logn, int rows, const uint8_t *seed, uint64_t stream, int dropped,
                              const uint64_t *psi, uint64_t inv_n, uint64_t *out, hipStream_t s) {
    switch (logn) {
        case 8: hipLaunchKernelGGL(k_expand_mask<8>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        case 11: hipLaunchKernelGGL(k_expand_mask<11>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_seat(int logn, int rows, const uint64_t *y, const uint64_t *psi, uint64_t inv_n, uint64_t *out,
                       hipStream_t s) {
    switch (logn) {
        case 8: hipLaunchKernelGGL(k_seat<8>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        case 11: hipLaunchKernelGGL(k_seat<11>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_store_build(const CtxDev &c, const uint16_t *records, uint16_t *col_scratch, const uint64_t *tag_key,
                              uint64_t *store, uint64_t *tags, hipStream_t s) {
    size_t total = (size_t)c.n * c.gamma;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, records, c.n, c.gamma,
                       col_scratch);
    switch (c.logn) {
        case 8: hipLaunchKernelGGL(k_store_build<8>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        case 11: hipLaunchKernelGGL(k_store_build<11>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_column_tags, dim3(c.gamma), dim3(kThreads), 0, s, col_scratch, tag_key, c.n, tags);
    return hipGetLastError();
}

}  // namespace inspire
