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
                                                           const uint64_t *psi, uint64_t inv_n, Pair *out) {
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
    for (int j = threadIdx.x; j < N; j += kThreads) {
        uint64_t v = mulmod(x[lds_ix(j)], inv_n);
        out[(size_t)blockIdx.x * N + j] = Pair{v, shoup_companion(v)};
    }
}

template <int LOGN>
__global__ __launch_bounds__(kThreads) void k_seat(const uint64_t *y, const uint64_t *psi, uint64_t inv_n,
                                                    Pair *out) {
    constexpr int N = 1 << LOGN;
    __shared__ uint64_t x[LdsSize<LOGN>::value];
    for (int j = threadIdx.x; j < N; j += kThreads) x[lds_ix(j)] = y[(size_t)blockIdx.x * N + j];
    __syncthreads();
    ntt_forward<LOGN>(x, psi);
    for (int j = threadIdx.x; j < N; j += kThreads) {
        uint64_t v = mulmod(x[lds_ix(j)], inv_n);
        out[(size_t)blockIdx.x * N + j] = Pair{v, shoup_companion(v)};
    }
}

// Folds the derived last digit into three rows: (r0, r1, r2) <- (r0 + k0 r2, r1 + k1 r2, kx r2).
__global__ void k_fold_rows(Pair *rows, int n, uint64_t kx, uint64_t k0, uint64_t k1) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint64_t r0 = rows[j].w, r1 = rows[n + j].w, r2 = rows[2 * n + j].w;
    uint64_t n0 = addmod(r0, mulmod(k0, r2)), n1 = addmod(r1, mulmod(k1, r2)), n2 = mulmod(kx, r2);
    rows[j] = Pair{n0, shoup_companion(n0)};
    rows[n + j] = Pair{n1, shoup_companion(n1)};
    rows[2 * n + j] = Pair{n2, shoup_companion(n2)};
}

hipError_t launch_fold_rows(Pair *rows, int n, uint64_t kx, uint64_t k0, uint64_t k1, hipStream_t s) {
    hipLaunchKernelGGL(k_fold_rows, dim3((n + 255) / 256), dim3(256), 0, s, rows, n, kx, k0, k1);
    return hipGetLastError();
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
__global__ __launch_bounds__(kThreads) void k_store_build(CtxDev c, const uint16_t *col, Pair *store) {
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
    for (int k = threadIdx.x; k < N; k += kThreads) {
        uint64_t v = mulmod(x[lds_ix(automorph_slot(t, k, LOGN))], c.inv_n_gamma);
        store[(size_t)i * N + k] = Pair{v, shoup_companion(v)};
    }
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

// DB scan, one lane per (row split, 4-column group): the lane walks its split's rows and
// keeps all 16 partial products of its four columns in registers, so there is no
// cross-lane reduction and no LDS. A wave's 64 lanes read one row's 64 consecutive u64
// words (512 contiguous bytes); the row's b and tag-key words are the same on every lane
// of the wave. Eight rows are in flight per lane.
// Partials: scan_part[q][split][column][4] = (sum h*b_lo, sum h*b_hi, sum h*key_lo, sum h*key_hi)
// over the split's rows, 32-bit halves in u64 (below 2^56 for n <= 4096); the chain kernel
// adds the splits. brev[k] is b laid out for coefficient 0 of b*h (reverse_for_coeff0,
// shard_store.rs:378-391); the sums are check_columns_and_fold_b's (shard_store.rs:197-265).
template <bool UNIFORM_ROW>
__global__ __launch_bounds__(kThreads) void k_scan(CtxDev c, BatchDev bt) {
    const int qi = blockIdx.x;
    const int n = c.n, gamma = c.gamma;
    const int G = gamma / 4;                     // 4-column groups per row (gamma >= 4)
    const int per_block = blockDim.x / G;        // splits per block
    const int g = threadIdx.x % G;
    int split = blockIdx.y * per_block + threadIdx.x / G;
    // with 64 or more groups a wave lies inside one split, so the row words are scalar
    if (UNIFORM_ROW) split = __builtin_amdgcn_readfirstlane(split);
    // ids of device-buffer batches are not checked on the host: a query naming no shard with
    // a store is skipped here and refused by k_chain (status -1)
    const ShardDev sh = shard_entry(c.shards, c.num_shards, bt.shard_ids[qi]);
    if (!shard_ready(sh)) return;
    const uint64_t *b = bt.b + (size_t)qi * n;
    const int rows = n / bt.ks, r0 = split * rows;  // a multiple of 8: n >= 256, ks <= n / 128
    const uint64_t *rec = reinterpret_cast<const uint64_t *>(sh.records) + g;
    uint64_t acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u][0] = acc[u][1] = acc[u][2] = acc[u][3] = 0;
    for (int k0 = r0; k0 < r0 + rows; k0 += 8) {
        uint64_t h4[8], bv[8], kv[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int k = k0 + r;
            h4[r] = rec[(size_t)k * G];
            bv[r] = b[(n - k) & (n - 1)];
            kv[r] = sh.tag_key[k];
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int k = k0 + r;
            // brev[0] = b[0], brev[k] = -b[n-k]
            const uint64_t bk = (k != 0 && bv[r] != 0) ? kQ - bv[r] : bv[r];
            const uint32_t bl = (uint32_t)bk, bh = (uint32_t)(bk >> 32);
            const uint32_t kl = (uint32_t)kv[r], kh = (uint32_t)(kv[r] >> 32);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t h = (uint32_t)(h4[r] >> (16 * u)) & 0xffffu;
                acc[u][0] += (uint64_t)h * bl;
                acc[u][1] += (uint64_t)h * bh;
                acc[u][2] += (uint64_t)h * kl;
                acc[u][3] += (uint64_t)h * kh;
            }
        }
    }
    uint64_t *o = bt.scan_part + (((size_t)qi * bt.ks + split) * gamma + 4 * g) * 4;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) o[u * 4 + v] = acc[u][v];
}

// Groups a launch's queries by shard (count <= kGroupMax, one workgroup of 1024 threads): a
// bitonic sort of the keys (shard id, query index) in LDS, then the scan tasks: each shard's run
// of queries in that order cut into tasks of two (the last of an odd run takes one). first[t]
// holds task t's start position | its length << 16, *ngroups the number of tasks.
__global__ __launch_bounds__(1024) void k_group(const uint32_t *ids, int count, uint32_t *order, uint32_t *first,
                                                uint32_t *ngroups) {
    constexpr int T = 1024, PER = kGroupMax / T;
    __shared__ uint64_t key[kGroupMax];
    __shared__ int part[T];
    const int t = threadIdx.x;
    int P = 1;
    while (P < count) P <<= 1;
    for (int i = t; i < P; i += T) key[i] = i < count ? ((uint64_t)ids[i] << 32 | (uint32_t)i) : ~0ULL;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < P; i += T) {
                const int l = i ^ j;
                if (l > i) {
                    const uint64_t a = key[i], b = key[l];
                    if ((a > b) == ((i & k) == 0)) {
                        key[i] = b;
                        key[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    // thread t owns positions PER t .. PER t + PER - 1; a run starts where the shard changes
    auto run_start = [&](int i) { return i == 0 || (key[i] >> 32) != (key[i - 1] >> 32); };
    int last = -1;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = t * PER + u;
        if (i < count) {
            order[i] = (uint32_t)key[i];
            if (run_start(i)) last = i;
        }
    }
    // inclusive max-scan of the last run start up to each thread's range
    part[t] = last;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        const int v = t >= off ? part[t - off] : -1;
        __syncthreads();
        if (v > part[t]) part[t] = v;
        __syncthreads();
    }
    int rs = t > 0 ? part[t - 1] : 0;  // start of the run the thread's first position is in
    __syncthreads();
    uint32_t task[PER];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = t * PER + u;
        task[u] = 0;
        if (i < count) {
            if (run_start(i)) rs = i;
            if (((i - rs) & 1) == 0) {
                const int len = i + 1 < count && !run_start(i + 1) ? 2 : 1;
                task[u] = (uint32_t)i | (uint32_t)len << 16 | 0x80000000u;
                cnt++;
            }
        }
    }
    part[t] = cnt;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {  // inclusive prefix sum of the per-thread task counts
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int ti = part[t] - cnt;
#pragma unroll
    for (int u = 0; u < PER; u++)
        if (task[u]) first[ti++] = task[u] & 0x7fffffffu;
    if (t == T - 1) *ngroups = (uint32_t)part[T - 1];
}

// One pass of the grouped scan over a lane's rows: the b fold of NQ queries and the column tag
// (the same for every query of the shard), accumulated in registers.
template <int NQ, int RF>
__device__ __forceinline__ void grouped_pass(const uint64_t *rec, int G, const uint64_t *tag_key,
                                             const uint64_t *b0, const uint64_t *b1, int n, int r0, int rows,
                                             uint64_t acc[4][2][2], uint64_t tag[4][2]) {
    const uint64_t *const b[2] = {b0, b1};
    for (int k0 = r0; k0 < r0 + rows; k0 += RF) {  // RF rows in flight
        uint64_t h4[RF], kv[RF], bv[NQ][RF];
#pragma unroll
        for (int r = 0; r < RF; r++) {
            const int k = k0 + r;
            h4[r] = rec[(size_t)k * G];
            kv[r] = tag_key[k];
#pragma unroll
            for (int v = 0; v < NQ; v++) bv[v][r] = b[v][(n - k) & (n - 1)];
        }
#pragma unroll
        for (int r = 0; r < RF; r++) {
            const int k = k0 + r;
            uint32_t bl[NQ], bh[NQ];
#pragma unroll
            for (int v = 0; v < NQ; v++) {
                // brev[0] = b[0], brev[k] = -b[n-k]
                const uint64_t bk = (k != 0 && bv[v][r] != 0) ? kQ - bv[v][r] : bv[v][r];
                bl[v] = (uint32_t)bk;
                bh[v] = (uint32_t)(bk >> 32);
            }
            const uint32_t kl = (uint32_t)kv[r], kh = (uint32_t)(kv[r] >> 32);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t h = (uint32_t)(h4[r] >> (16 * u)) & 0xffffu;
#pragma unroll
                for (int v = 0; v < NQ; v++) {
                    acc[u][v][0] += (uint64_t)h * bl[v];
                    acc[u][v][1] += (uint64_t)h * bh[v];
                }
                tag[u][0] += (uint64_t)h * kl;
                tag[u][1] += (uint64_t)h * kh;
            }
        }
    }
}

// DB scan of a grouped launch: one workgroup per (task, block of row splits), lanes laid out as
// in k_scan. A task is one or two queries of one shard (k_group), so each record word is read
// from HBM once per task and folded against the b of both queries, and the column tag, the same
// for every query of the shard, is summed once per task. The workgroup then adds its row splits'
// partials in LDS and writes one slice per query (scan_slices), a quarter of k_scan's partials
// at ring 2048, so the packing kernel reads fewer. Its per-query tag check is unchanged.
template <bool UNIFORM_ROW>
__device__ __forceinline__ void scan_grouped_body(const CtxDev &c, const BatchDev &bt) {
    // per lane: 4 columns x 2 halves of the tag sums, then of one query's fold sums (16 KiB)
    __shared__ uint64_t red[kThreads * 8];
    const int task = blockIdx.x;
    if (task >= (int)*bt.ngroups) return;
    const int n = c.n, gamma = c.gamma;
    const int G = gamma / 4;
    const int threads = blockDim.x;
    const int per_block = threads / G;
    const int g = threadIdx.x % G, w = threadIdx.x / G;
    int split = blockIdx.y * per_block + w;
    if (UNIFORM_ROW) split = __builtin_amdgcn_readfirstlane(split);
    // the task's queries are the same on every lane: scalar registers
    const uint32_t tw = __builtin_amdgcn_readfirstlane(bt.first[task]);
    const int p0 = (int)(tw & 0xffffu), nq = (int)(tw >> 16);
    uint32_t qi[2];
    qi[0] = __builtin_amdgcn_readfirstlane(bt.order[p0]);
    qi[1] = __builtin_amdgcn_readfirstlane(bt.order[nq == 2 ? p0 + 1 : p0]);
    // (indices made scalar, pointers derived from the kernel arguments: so the loads of b and of
    // the tag key stay scalar loads)
    const uint32_t sid = __builtin_amdgcn_readfirstlane(bt.shard_ids[qi[0]]);
    const ShardDev sh = shard_entry(c.shards, c.num_shards, sid);
    if (!shard_ready(sh)) return;
    const int rows = n / bt.ks, r0 = split * rows;
    // the record and tag-key pointers stay generic here (glob, device.hpp, is not applied): as
    // global pointers the uniform tag-key loads become scalar loads, and the scan then ran 20%
    // slower on the 2^20-record launch (0.109 against 0.088 ms) than with vector loads of them
    const uint64_t *rec = reinterpret_cast<const uint64_t *>(sh.records) + g;
    const uint64_t *b0 = bt.b + (size_t)qi[0] * n, *b1 = bt.b + (size_t)qi[1] * n;
    uint64_t acc[4][2][2], tag[4][2];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        tag[u][0] = tag[u][1] = 0;
        acc[u][0][0] = acc[u][0][1] = acc[u][1][0] = acc[u][1][1] = 0;
    }
    if (nq == 2)
        grouped_pass<2, 2>(rec, G, sh.tag_key, b0, b1, n, r0, rows, acc, tag);
    else
        grouped_pass<1, 8>(rec, G, sh.tag_key, b0, b1, n, r0, rows, acc, tag);
    // add the workgroup's row splits in LDS and write one slice per query: the tag sums once
    // (the same for both queries), then each query's fold sums. A slice holds, per column j,
    // (fold lo, fold hi, tag lo, tag hi) at 4 j.
    const int slice = blockIdx.y, outs = G * 8;
    uint64_t *mine = red + (size_t)(w * G + g) * 8;
    uint64_t *o0 = bt.scan_part + ((size_t)qi[0] * bt.ks + slice) * gamma * 4;
    uint64_t *o1 = bt.scan_part + ((size_t)qi[1] * bt.ks + slice) * gamma * 4;
    for (int round = 0; round <= nq; round++) {  // round 0: tags, round 1 + v: query v's folds
#pragma unroll
        for (int u = 0; u < 4; u++) {
            mine[2 * u + 0] = round == 0 ? tag[u][0] : round == 1 ? acc[u][0][0] : acc[u][1][0];
            mine[2 * u + 1] = round == 0 ? tag[u][1] : round == 1 ? acc[u][0][1] : acc[u][1][1];
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < outs; idx += threads) {
            uint64_t sum = 0;
            for (int ww = 0; ww < per_block; ww++) sum += red[(size_t)ww * G * 8 + idx];
            // idx = 8 g' + 2 u + e: column 4 g' + u, entry e of the fold or 2 + e of the tag
            const int at = (idx >> 1) * 4 + (idx & 1);
            if (round == 0) {
                o0[at + 2] = sum;
                if (nq == 2) o1[at + 2] = sum;
            } else {
                (round == 1 ? o0 : o1)[at] = sum;
            }
        }
        __syncthreads();
    }
}

// Wide shards (64 or more column groups, the row words scalar): registers budgeted for six
// waves per SIMD, so a 2^20-record launch's workgroups run in one round.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8)))
void k_scan_grouped(CtxDev c, BatchDev bt) { scan_grouped_body<true>(c, bt); }
__global__ __launch_bounds__(kThreads) void k_scan_grouped_narrow(CtxDev c, BatchDev bt) {
    scan_grouped_body<false>(c, bt);
}

// Narrow shards (gamma < 4 columns): one column per lane.
__global__ __launch_bounds__(kThreads) void k_scan_narrow(CtxDev c, BatchDev bt) {
    const int qi = blockIdx.x, split = blockIdx.y;
    const int n = c.n, gamma = c.gamma;
    // ids of device-buffer batches are not checked on the host: a query naming no shard with
    // a store is skipped here and refused by k_chain (status -1)
    const ShardDev sh = shard_entry(c.shards, c.num_shards, bt.shard_ids[qi]);
    if (!shard_ready(sh)) return;
    const uint64_t *b = bt.b + (size_t)qi * n;
    const int rows = n / bt.ks, r0 = split * rows;
    const int lanes = kThreads / gamma;
    const int j = threadIdx.x % gamma, lane = threadIdx.x / gamma;
    uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int k = r0 + lane; k < r0 + rows; k += lanes) {
        uint64_t h = sh.records[(size_t)k * gamma + j];
        uint64_t bk = k == 0 ? b[0] : (b[n - k] ? kQ - b[n - k] : 0);
        uint64_t key = sh.tag_key[k];
        a0 += h * (bk & 0xffffffffu);
        a1 += h * (bk >> 32);
        a2 += h * (key & 0xffffffffu);
        a3 += h * (key >> 32);
    }
    __shared__ uint64_t red[kThreads][4];
    red[threadIdx.x][0] = a0; red[threadIdx.x][1] = a1; red[threadIdx.x][2] = a2; red[threadIdx.x][3] = a3;
    __syncthreads();
    if ((int)threadIdx.x < gamma) {
        uint64_t s[4] = {0, 0, 0, 0};
        for (int l = 0; l < lanes; l++)
            for (int v = 0; v < 4; v++) s[v] += red[l * gamma + threadIdx.x][v];
        uint64_t *o = bt.scan_part + (((size_t)qi * bt.ks + split) * gamma + threadIdx.x) * 4;
        for (int v = 0; v < 4; v++) o[v] = s[v];
    }
}

// ---------------------------------------------------------------------------
// launchers


hipError_t launch_group_only(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    if (b.count == 0 || !batch_grouped(c, b)) return hipSuccess;
    hipLaunchKernelGGL(k_group, dim3(1), dim3(1024), 0, s, b.shard_ids, b.count, b.order, b.first, b.ngroups);
    return hipGetLastError();
}

// The DB scan of a launch whose grouping (launch_group_only) is already enqueued.
hipError_t launch_scan_only(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    if (b.count == 0) return hipSuccess;
    if (batch_grouped(c, b)) {
        const int G = c.gamma / 4;
        const int threads = scan_threads(c.gamma, b.ks);
        // one workgroup row per possible task (at most one per query); the tasks past the count
        // k_group found exit at once
        if (G >= 64)
            hipLaunchKernelGGL(k_scan_grouped, dim3(b.count, b.ks / (threads / G)), dim3(threads), 0, s, c, b);
        else
            hipLaunchKernelGGL(k_scan_grouped_narrow, dim3(b.count, b.ks / (threads / G)), dim3(threads), 0, s, c,
                               b);
        return hipGetLastError();
    }
    if (c.gamma >= 4) {
        // one lane per (split, group): a block holds whole splits, the grid the rest
        const int G = c.gamma / 4;
        const int threads = G * b.ks < kThreads ? G * b.ks : kThreads;
        if (G >= 64)
            hipLaunchKernelGGL(k_scan<true>, dim3(b.count, b.ks / (threads / G)), dim3(threads), 0, s, c, b);
        else
            hipLaunchKernelGGL(k_scan<false>, dim3(b.count, b.ks / (threads / G)), dim3(threads), 0, s, c, b);
    }
    else
        hipLaunchKernelGGL(k_scan_narrow, dim3(b.count, b.ks), dim3(kThreads), 0, s, c, b);
    return hipGetLastError();
}

hipError_t launch_respond(const CtxDev &c, const BatchDev &b, hipStream_t s) {
    if (b.count == 0) return hipSuccess;
    hipError_t e = launch_group_only(c, b, s);
    if (e == hipSuccess) e = launch_scan_only(c, b, s);
    if (e != hipSuccess) return e;
    return launch_chain_only(c, b, s);
}


hipError_t launch_expand_mask(int logn, int rows, const uint8_t *seed, uint64_t stream, int dropped,
                              const uint64_t *psi, uint64_t inv_n, Pair *out, hipStream_t s) {
    switch (logn) {
        case 8: hipLaunchKernelGGL(k_expand_mask<8>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        case 9: hipLaunchKernelGGL(k_expand_mask<9>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        case 10: hipLaunchKernelGGL(k_expand_mask<10>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        case 11: hipLaunchKernelGGL(k_expand_mask<11>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        case 12: hipLaunchKernelGGL(k_expand_mask<12>, dim3(rows), dim3(kThreads), 0, s, seed, stream, dropped, psi, inv_n, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_seat(int logn, int rows, const uint64_t *y, const uint64_t *psi, uint64_t inv_n, Pair *out,
                       hipStream_t s) {
    switch (logn) {
        case 8: hipLaunchKernelGGL(k_seat<8>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        case 9: hipLaunchKernelGGL(k_seat<9>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        case 10: hipLaunchKernelGGL(k_seat<10>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        case 11: hipLaunchKernelGGL(k_seat<11>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        case 12: hipLaunchKernelGGL(k_seat<12>, dim3(rows), dim3(kThreads), 0, s, y, psi, inv_n, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_store_build(const CtxDev &c, const uint16_t *records, uint16_t *col_scratch, const uint64_t *tag_key,
                              Pair *store, uint64_t *tags, hipStream_t s) {
    size_t total = (size_t)c.n * c.gamma;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, records, c.n, c.gamma,
                       col_scratch);
    switch (c.logn) {
        case 8: hipLaunchKernelGGL(k_store_build<8>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        case 9: hipLaunchKernelGGL(k_store_build<9>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        case 10: hipLaunchKernelGGL(k_store_build<10>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        case 11: hipLaunchKernelGGL(k_store_build<11>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        case 12: hipLaunchKernelGGL(k_store_build<12>, dim3(c.gamma), dim3(kThreads), 0, s, c, col_scratch, store); break;
        default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_column_tags, dim3(c.gamma), dim3(kThreads), 0, s, col_scratch, tag_key, c.n, tags);
    return hipGetLastError();
}

}  // namespace inspire
