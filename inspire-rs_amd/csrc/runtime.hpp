// Internal runtime objects shared by the C ABI (capi.cpp) and the serving engine
// (server.cpp). Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/inspire_gpu.h"
#include "device.hpp"
#include "fence.hpp"
#include "kernels.hpp"
#include "modq.hpp"
#include "trace.hpp"

using namespace inspire;
typedef unsigned __int128 u128;

namespace inspire_rt {

inline void set_err(char *err, size_t cap, const char *fmt, ...) {
    if (!err || cap == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, cap, fmt, ap);
    va_end(ap);
}

inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    while (e) {
        if (e & 1) r = (uint64_t)((u128)r * b % m);
        b = (uint64_t)((u128)b * b % m);
        e >>= 1;
    }
    return r;
}
inline uint64_t inv_mod(uint64_t a) { return powmod(a, kQ - 2, kQ); }

// psi: a primitive 2n-th root of unity, g^((q-1)/2n) for the least quadratic
// non-residue g (NttContext::find_primitive_root, src/math/ntt.rs:743-752).
inline uint64_t find_psi(uint64_t two_n) {
    uint64_t ex = (kQ - 1) / two_n;
    for (uint64_t g = 2; g < 65536; g++) {
        uint64_t c = powmod(g, ex, kQ);
        if (powmod(c, two_n, kQ) == 1 && powmod(c, two_n / 2, kQ) != 1) return c;
    }
    return 0;
}

// ladder[m] = psi^brv(m) built as the reference does (ntt.rs:797-834), standard form.
inline std::vector<uint64_t> ladder(size_t n, uint64_t psi) {
    std::vector<uint64_t> f(n, 0);
    if (n > 1) f[1] = 1;
    for (size_t m = 1; m < n; m++) {
        if ((m & (m - 1)) == 0)
            f[m] = powmod(psi, n / (2 * m), kQ);
        else
            f[m] = (uint64_t)((u128)f[m & (m - 1)] * f[m & (~m + 1)] % kQ);
    }
    return f;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = b;
        return hipMalloc(&p, b ? b : 8);
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

inline bool os_random(void *out, size_t len) {
    std::ifstream f("/dev/urandom", std::ios::binary);
    if (!f) return false;
    f.read(static_cast<char *>(out), (std::streamsize)len);
    return (size_t)f.gcount() == len;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            set_err(err, err_cap, "HIP failure in %s: %s", #expr, hipGetErrorString(e_)); \
            return INSPIRE_ERR_DEVICE;                                                 \
        }                                                                              \
    } while (0)

}  // namespace inspire_rt

using namespace inspire_rt;

struct inspire_shard {
    inspire_ctx *ctx = nullptr;
    uint32_t id = 0;
    DevBuf records, store, tag_key, tags;
    bool has_store = false;
    bool lazy_store = false;  // built by a respond, not by precompute_packing
};

struct inspire_seat {
    inspire_ctx *ctx = nullptr;
    uint32_t rows = 0, num_to_pack = 0;
    DevBuf yprime;
    // recorded on the context's stream behind the kernels that fill yprime; a launch on another
    // stream waits for it until it has been seen complete (settled)
    hipEvent_t ready = nullptr;
    bool settled = false;
    ~inspire_seat() {
        if (ready) (void)hipEventDestroy(ready);
    }
};

struct inspire_workspace {
    inspire_ctx *ctx = nullptr;
    size_t max_count = 0;
    DevBuf scan_part, status;
    DevBuf group;  // shard grouping of a launch: order, first, ngroups (up to kGroupMax queries)
    // staging for the host-buffer entry point
    DevBuf ids, seeds, a_in, b, a_out, b_out, wire;
    size_t wire_cap = 0;
    KernelTrace trace;  // scan and chain events of the next device-buffer launches, when armed
};

// Refusal of a respond or server over a context created without a CRS.
constexpr const char *kNoCrs =
    "context has no CRS: it was created without a w seed (shards and packing stores only); bind the CRS with "
    "inspire_ctx_attach_crs before responding";

struct inspire_ctx {
    std::mutex mu;
    int logn = 0, n = 0, gamma = 0, rows = 0, dropped = 0;
    uint8_t w_seed[32];
    std::atomic<bool> has_crs{false};  // wprime derived from w_seed (inspire_ctx_attach_crs)
    DevBuf psi, psi_inv, twf, twi, wprime, gpow, gpow_back, shard_table, col_scratch;
    uint64_t inv_n = 0, inv_n_gamma = 0;
    Pair kx{0, 0}, k0{0, 0}, k1{0, 0};
    std::vector<ShardDev> shards;        // dense by id
    std::vector<inspire_shard *> owners;
    bool table_dirty = true;
    int queries_per_cu = 1;
    int device = 0;  // the GPU the context's memory lives on (current at creation)
    inspire_workspace *default_ws = nullptr;
    std::vector<void *> retired;  // superseded shard tables
    // the host-buffer entry points and store builds launch here and wait for this stream
    // alone, not for the whole device (other callers' streams keep running)
    hipStream_t stream = nullptr;
    // every launch that reads the shard table, a shard, a seat or a workspace records an event
    // here under mu, so destroying one of them waits for those launches only
    LaunchFence fence;
    // seats' device memory (rows x n Pairs each), reused across registrations; the host rows of
    // a seat being created are staged in seat_staging (rows x n words), ordered on stream
    BlockPool seat_pool;
    DevBuf seat_staging;
    ~inspire_ctx() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        (void)wait_for(fence.snapshot());
        for (void *p : retired) (void)hipFree(p);
    }

    // Row splits of the DB scan at ring 2048 and up: 16 unless inspire_ctx_autotune_scan picked
    // another of kScanSplitChoices. Scan scratch is sized for kMaxScanSplits, so a change applies
    // to the next launch of every workspace and serving slot; each launch reads it once.
    std::atomic<int> scan_splits{16};
    int ks() const { return n >= 2048 ? scan_splits.load(std::memory_order_relaxed) : 1; }
    int ks_cap() const { return n >= 2048 ? kMaxScanSplits : 1; }

    CtxDev dev() const {
        CtxDev c;
        c.logn = logn; c.n = n; c.gamma = gamma; c.rows = rows; c.exact = dropped == 0;
        c.psi = psi.as<uint64_t>(); c.psi_inv = psi_inv.as<uint64_t>();
        c.twf = twf.as<Pair>(); c.twi = twi.as<Pair>();
        c.wprime = wprime.as<Pair>();
        c.gpow = gpow.as<uint32_t>(); c.gpow_back = gpow_back.as<uint32_t>();
        c.inv_n_gamma = inv_n_gamma; c.kx = kx; c.k0 = k0; c.k1 = k1;
        c.shards = shard_table.as<ShardDev>();
        c.num_shards = (int)shards.size();
        c.queries_per_cu = queries_per_cu;
        return c;
    }
};

namespace inspire_rt {

inline int sync_table(inspire_ctx *ctx, char *err, size_t err_cap) {
    if (!ctx->table_dirty) return INSPIRE_OK;
    size_t bytes = ctx->shards.size() * sizeof(ShardDev);
    if (ctx->shard_table.bytes < bytes || !ctx->shard_table.p) {
        // a launch in flight may still read the old table: retire it, free it with the context
        if (ctx->shard_table.p) {
            ctx->retired.push_back(ctx->shard_table.p);
            ctx->shard_table.p = nullptr;
        }
        HIP_TRY(ctx->shard_table.alloc(bytes * 2 + 64 * sizeof(ShardDev)));
    }
    if (bytes) HIP_TRY(hipMemcpy(ctx->shard_table.p, ctx->shards.data(), bytes, hipMemcpyHostToDevice));
    ctx->table_dirty = false;
    return INSPIRE_OK;
}

// Builds a shard's packing store on the device (capi.cpp); caller holds ctx->mu. lazy marks a
// store built by a respond for a shard that was uploaded without one: the reference packs such
// a shard from its columns on every query, so an edit of the shard drops this store (the next
// respond builds it again from the edited columns) instead of failing its tag check.
int build_store(inspire_ctx *ctx, inspire_shard *sh, bool lazy, char *err, size_t err_cap);

// inspire_seat_create without the wait: the seat's kernels are enqueued on the context's stream
// and seat->ready recorded behind them (capi.cpp).
int seat_create_enqueue(inspire_ctx *ctx, const uint64_t *y_body, uint32_t rows, uint32_t num_to_pack,
                        inspire_seat **out, char *err, size_t err_cap);

inline int ws_reserve(inspire_workspace *ws, size_t count, char *err, size_t err_cap) {
    inspire_ctx *ctx = ws->ctx;
    if (count <= ws->max_count && ws->scan_part.p) return INSPIRE_OK;
    if (count > INSPIRE_MAX_BATCH) {
        set_err(err, err_cap, "a workspace holds at most %u queries, asked for %zu", INSPIRE_MAX_BATCH, count);
        return INSPIRE_ERR_INVALID;
    }
    size_t c = count < 1 ? 1 : count;
    HIP_TRY(ws->scan_part.alloc(c * ctx->ks_cap() * ctx->gamma * 4 * sizeof(uint64_t)));
    HIP_TRY(ws->status.alloc(c * sizeof(int32_t)));
    const size_t gc = c < (size_t)kGroupMax ? c : (size_t)kGroupMax;
    HIP_TRY(ws->group.alloc((2 * gc + 2) * sizeof(uint32_t)));
    ws->max_count = c;
    return INSPIRE_OK;
}

// Points a launch of count queries at the workspace's grouping buffers when it can be grouped.
inline void ws_group(const inspire_workspace *ws, BatchDev &b) {
    const size_t gc = ws->max_count < (size_t)kGroupMax ? ws->max_count : (size_t)kGroupMax;
    if ((size_t)b.count > gc || !ws->group.p) return;
    uint32_t *g = ws->group.as<uint32_t>();
    b.order = g;
    b.first = g + gc;
    b.ngroups = g + 2 * gc + 1;
}

}  // namespace inspire_rt
