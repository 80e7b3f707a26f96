// Device-side descriptors shared by the kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xcd.hpp"

namespace inspire {

// A fixed multiplicand w < q with its Shoup companion wp = floor(w * 2^64 / q).
struct Pair {
    uint64_t w, wp;
};

// One shard resident in HBM (the reference's ShardData + ShardPackingStore,
// src/pir/setup.rs:280-289, src/inspiring/shard_store.rs:36-48).
struct ShardDev {
    const uint16_t *records;  // ring_dim x gamma, row-major: column j of record m at [m*gamma + j]
    const Pair *store;        // gamma x ring_dim: n^-1 * S_i in NTT form, with companions
    const uint64_t *tag_key;  // ring_dim tag key words, each < 2^61 - 1
    const uint64_t *tags;     // gamma column tags mod 2^61 - 1
};

// The shard entry a query names, read once: zero when the id is past the table, and usable only
// when every pointer is set (a shard destroyed, or its store still being built, while a launch
// naming it runs is seen with some or all of them null, never half-freed: destroying a shard
// waits for the launches enqueued before it).
__device__ inline ShardDev shard_entry(const ShardDev *table, int num_shards, uint32_t sid) {
    ShardDev z{nullptr, nullptr, nullptr, nullptr};
    return sid < (uint32_t)num_shards ? table[sid] : z;
}
__device__ inline bool shard_ready(const ShardDev &s) { return s.records && s.store && s.tag_key && s.tags; }
// Device memory through the global address space. A pointer read from memory (the shard table,
// a seat array, a job entry) is generic to the compiler, and a generic (flat) load also counts
// against the LDS and scalar-load counter, so every LDS or scalar wait behind it waits for HBM as
// well, and a uniform generic load cannot become a scalar load. Casting such pointers keeps the
// loads global.
template <class T>
using Glob = const __attribute__((address_space(1))) T *;
template <class T>
__device__ inline Glob<T> glob(const T *p) { return (Glob<T>)p; }
// A pointer every lane holds alike, kept in scalar registers.
template <class T>
__device__ inline T *uniform(T *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T *>((uint64_t)hi << 32 | lo);
}

// Per-CRS constants (the reference's PackParams + OfflinePackingKeys,
// src/inspiring/inspiring2.rs:28-60, 450-470), all on the device.
struct CtxDev {
    int logn, n, gamma, rows, exact;
    const uint64_t *psi;        // n, forward ladder psi^brv(m), standard form
    const uint64_t *psi_inv;    // n, inverse ladder
    const Pair *twf;            // n, forward ladder with companions
    const Pair *twi;            // n, inverse ladder with companions
    const Pair *wprime;         // rows x n: n^-1 * NTT(w_mask_k)
    const uint32_t *gpow;       // gamma: g^i mod 2n
    const uint32_t *gpow_back;  // gamma: g^(n-i mod n) mod 2n
    uint64_t inv_n_gamma;       // (n * gamma)^-1 mod q
    Pair kx, k0, k1;            // last-digit constants (exact gadget)
    const ShardDev *shards;     // dense shard table
    int num_shards;
    int queries_per_cu;         // launch shaping: 1 spreads queries over CUs, 2 co-schedules pairs
};

// Inputs and outputs of one respond batch, device pointers.
struct BatchDev {
    int count;
    const uint32_t *shard_ids;  // count
    const uint8_t *seeds;       // count x 32 (seeded queries), or
    const uint64_t *a_in;       // count x n expanded a (unseeded queries) when non-null
    const uint64_t *b;          // count x n
    const Pair *yprime;         // rows x n: n^-1 * NTT(y_body_k) of the seat, or
    const Pair *const *yprimes; // per-query seats (count pointers) when non-null
    uint64_t *scan_part;        // count x ks x gamma x 4 scratch
    int ks;                     // row splits of the scan
    uint64_t *a_out;            // count x n
    uint64_t *b_out;            // count x gamma
    int32_t *status;            // count: 0 ok, c+1 = column c failed its tag, -1 = no shard with a store
    uint64_t switch_modulus;    // 0 = no modulus switch
    // Shard grouping of the launch (a workspace's or a serving slot's scratch, count <= kGroupMax;
    // all null for an ungrouped launch): k_group writes the query indices sorted by shard id
    // (order, count words), the start of each shard's run in that order (first, count + 1
    // words) and the number of runs (ngroups). The DB scan then reads each distinct shard once
    // for all of its queries, and the packing kernel takes its queries in that order, a
    // contiguous slice of it per XCD, so chains of one shard share their store rows in one L2.
    uint32_t *order = nullptr, *first = nullptr, *ngroups = nullptr;
    // when non-null, each query's ServerResponse bincode bytes at wire_out + qi * wire_stride
    // (4-byte aligned, stride a multiple of 4 holding the length rounded up to 4); a_out and
    // b_out may then be null
    uint8_t *wire_out = nullptr;
    size_t wire_stride = 0;
};

// Largest launch the grouping kernel sorts (one workgroup, keys in LDS); larger launches scan
// query by query.
constexpr int kGroupMax = 4096;
// Whether a launch is grouped by shard: it carries the group buffers, fits the sort, and its
// shards are wide enough for the grouped scan's lane layout. The scan and packing kernels both
// decide by this.
__host__ __device__ inline bool batch_grouped(const CtxDev &c, const BatchDev &b) {
    return b.order && b.first && b.ngroups && b.count <= kGroupMax && c.gamma >= 4;
}

// Row-split counts of the DB scan at ring 2048 and up (a launch's scan tile, picked by the
// autotuner; 16 by default): every split is a multiple of 8 rows (n / 32 = 64 at n = 2048) and a
// multiple of the four splits a wide shard's workgroup holds.
constexpr int kScanSplitChoices[3] = {8, 16, 32};
constexpr int kMaxScanSplits = 32;

// Workgroup size of the DB scans over 4-column groups (gamma >= 4): one lane per (row split,
// group), at most 256.
__host__ __device__ inline int scan_threads(int gamma, int ks) {
    const int G = gamma / 4;
    return G * ks < 256 ? G * ks : 256;
}
// Partial slices per query the scan leaves in scan_part (each query's stride stays ks slices):
// ks for a query-by-query scan; one per workgroup of row splits for a grouped one, which adds
// its splits in LDS.
__host__ __device__ inline int scan_slices(const CtxDev &c, const BatchDev &b) {
    if (!batch_grouped(c, b)) return b.ks;
    return b.ks / (scan_threads(c.gamma, b.ks) / (c.gamma / 4));
}

// Layout of a packed InspiRING ServerResponse in bincode (respond.rs:111-176, poly.rs:166-225,
// 286-310): u32 variant 1 | u64 A | a tight (A bytes) | u64 1 | u64 modulus | u64 modulus |
// u64 n | u64 0 | u8 0 | u64 B | b prefix tight (B bytes) | u32 gamma | u64 0 | u8 1 | u32 0.
struct WireLayout {
    uint32_t w, a_bytes, b_bytes, total, n, gamma;
    uint64_t modulus;
};

__host__ __device__ inline WireLayout wire_layout(uint32_t n, uint32_t gamma, uint64_t modulus) {
    WireLayout L;
    uint32_t w = 0;
    while (w < 64 && ((modulus - 1) >> w) != 0) w++;
    L.w = w;
    L.a_bytes = (n * w + 7) / 8;
    L.b_bytes = (gamma * w + 7) / 8;
    L.total = L.a_bytes + L.b_bytes + 78;
    L.n = n;
    L.gamma = gamma;
    L.modulus = modulus;
    return L;
}

}  // namespace inspire
