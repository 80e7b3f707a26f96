// Host runtime behind include/inspire_gpu.h: parameter checks, HBM-resident objects,
// and the launches of the respond pipeline.
#include "runtime.hpp"

namespace inspire_rt {

int build_store(inspire_ctx *ctx, inspire_shard *sh, bool lazy, char *err, size_t err_cap) {
    const size_t n = ctx->n, g = ctx->gamma;
    // key words below 2^61 - 1, drawn as the reference draws its tag key
    // (shard_store.rs:127-135): 61 random bits, rejected at or above the prime
    std::vector<uint64_t> key(n), draw(n);
    if (!os_random(draw.data(), n * 8)) {
        set_err(err, err_cap, "OS entropy unavailable for ShardPackingStore::build");
        return INSPIRE_ERR_DEVICE;
    }
    for (size_t k = 0; k < n; k++) {
        uint64_t v = draw[k] >> 3;
        while (v >= kTagPrime) {
            if (!os_random(&v, 8)) {
                set_err(err, err_cap, "OS entropy unavailable for ShardPackingStore::build");
                return INSPIRE_ERR_DEVICE;
            }
            v >>= 3;
        }
        key[k] = v;
    }
    if (!sh->store.p) HIP_TRY(sh->store.alloc(g * n * sizeof(Pair)));
    if (!sh->tag_key.p) HIP_TRY(sh->tag_key.alloc(n * 8));
    if (!sh->tags.p) HIP_TRY(sh->tags.alloc(g * 8));
    if (!ctx->col_scratch.p) HIP_TRY(ctx->col_scratch.alloc(g * n * 2));
    HIP_TRY(hipMemcpy(sh->tag_key.p, key.data(), n * 8, hipMemcpyHostToDevice));
    int rc = sync_table(ctx, err, err_cap);
    if (rc) return rc;
    HIP_TRY(launch_store_build(ctx->dev(), sh->records.as<uint16_t>(), ctx->col_scratch.as<uint16_t>(),
                               sh->tag_key.as<uint64_t>(), sh->store.as<Pair>(), sh->tags.as<uint64_t>(),
                               ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    sh->has_store = true;
    sh->lazy_store = lazy;
    ShardDev &d = ctx->shards[sh->id];
    d.store = sh->store.as<Pair>();
    d.tag_key = sh->tag_key.as<uint64_t>();
    d.tags = sh->tags.as<uint64_t>();
    ctx->table_dirty = true;
    return sync_table(ctx, err, err_cap);
}

}  // namespace inspire_rt

namespace {

// --- scratch of the allocation-free entry point (inspire_respond_batch_device_noalloc)
// The DB scan's partial sums (and the parked b_poly) of one launch live in this file-scope
// device array instead of a workspace: every GPU the library runs on holds its own copy, of
// INSPIRE_NOALLOC_SCRATCH_BYTES, until the process ends. In host code `scratch` names a
// host-side stand-in, not device memory: kernels and copies get the per-device address that
// hipGetSymbolAddress returns, looked up once per device when a context is created there.
static __device__ __align__(256) uint64_t scratch[INSPIRE_NOALLOC_SCRATCH_BYTES / 8];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// Caller has `device` current.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = ptr;
    return hipSuccess;
}

void *scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}

int register_shard(inspire_ctx *ctx, uint32_t id, const std::vector<uint16_t> &rec, inspire_shard **out, char *err,
                   size_t err_cap) {
    if (id < ctx->owners.size() && ctx->owners[id]) {
        set_err(err, err_cap, "shard %u is already resident; destroy it before uploading another at that position", id);
        return INSPIRE_ERR_INVALID;
    }
    inspire_shard *sh = new (std::nothrow) inspire_shard();
    if (!sh) { set_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    sh->ctx = ctx;
    sh->id = id;
    hipError_t e = sh->records.alloc(rec.size() * 2);
    if (e == hipSuccess) e = hipMemcpy(sh->records.p, rec.data(), rec.size() * 2, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_err(err, err_cap, "HIP failure uploading shard %u: %s", id, hipGetErrorString(e));
        delete sh;
        return INSPIRE_ERR_DEVICE;
    }
    if (ctx->shards.size() <= id) {
        ctx->shards.resize(id + 1, ShardDev{nullptr, nullptr, nullptr, nullptr});
        ctx->owners.resize(id + 1, nullptr);
    }
    ctx->shards[id] = ShardDev{sh->records.as<uint16_t>(), nullptr, nullptr, nullptr};
    ctx->owners[id] = sh;
    ctx->table_dirty = true;
    *out = sh;
    return INSPIRE_OK;
}

// Host-side refusals, before any launch (the reference checks every query's shape
// and shard before packing: respond.rs:409-611).
int check_batch(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *ids, size_t count, char *err,
                size_t err_cap) {
    if (!ctx->has_crs) {
        set_err(err, err_cap, kNoCrs);
        return INSPIRE_ERR_INVALID;
    }
    if (!seat || seat->ctx != ctx) {
        set_err(err, err_cap, "packing-key seat was built for another context");
        return INSPIRE_ERR_INVALID;
    }
    if (seat->num_to_pack != (uint32_t)ctx->gamma) {
        set_err(err, err_cap,
                "packing key was built for gamma = %u but the server packs gamma = %d; the client must re-derive "
                "its packing keys against the current CRS",
                seat->num_to_pack, ctx->gamma);
        return INSPIRE_ERR_INVALID;
    }
    for (size_t k = 0; ids && k < count; k++) {
        uint32_t id = ids[k];
        if (id >= ctx->owners.size() || !ctx->owners[id]) {
            set_err(err, err_cap,
                    "respond_seeded_with_seat: shard %u not found at dense position %u; encoded database has %zu "
                    "shards",
                    id, id, ctx->owners.size());
            return INSPIRE_ERR_INVALID;
        }
    }
    return INSPIRE_OK;
}

}  // namespace

int inspire_rt::seat_create_enqueue(inspire_ctx *ctx, const uint64_t *y_body, uint32_t rows, uint32_t num_to_pack,
                                     inspire_seat **out, char *err, size_t err_cap) {
    if (!ctx || !y_body || !out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    const size_t n = ctx->n;  // the context's shape never changes: checked without its lock
    if (num_to_pack != (uint32_t)ctx->gamma) {
        set_err(err, err_cap,
                "packing key was built for gamma = %u but the server packs gamma = %d; the client must re-derive its "
                "packing keys against the current CRS",
                num_to_pack, ctx->gamma);
        return INSPIRE_ERR_INVALID;
    }
    if (rows != (uint32_t)ctx->rows) {
        set_err(err, err_cap,
                "packing-key seat is for another ring or gadget: it holds %u key rows on ring_dim %zu, the server packs "
                "%d rows on ring_dim %zu. The client must re-derive its packing keys against the current CRS.",
                rows, n, ctx->rows, n);
        return INSPIRE_ERR_INVALID;
    }
    for (size_t i = 0; i < rows * n; i++)
        if (y_body[i] >= kQ) {
            set_err(err, err_cap, "Poly decode refused: limb 0 coefficient %zu is %llu, not canonical for modulus %llu",
                    i % n, (unsigned long long)y_body[i], (unsigned long long)kQ);
            return INSPIRE_ERR_INVALID;
        }
    inspire_seat *seat = new (std::nothrow) inspire_seat();
    if (!seat) { set_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    seat->ctx = ctx;
    seat->rows = rows;
    seat->num_to_pack = num_to_pack;
    // Enqueued on the context's stream under the lock: a registration while launches run
    // allocates and frees nothing that waits for the device (its memory comes from the seat
    // pool, the rows go through the context's staging buffer, which the next registration
    // overwrites only behind these kernels on the same stream), and nothing waits here.
    hipError_t e = hipEventCreateWithFlags(&seat->ready, hipEventDisableTiming);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->seat_pool.block = (size_t)rows * n * sizeof(Pair);
        void *mem = nullptr;
        if (e == hipSuccess) e = ctx->seat_pool.take(&mem);
        if (e == hipSuccess) {
            seat->yprime.p = mem;
            seat->yprime.bytes = ctx->seat_pool.block;
        }
        if (e == hipSuccess && !ctx->seat_staging.p) e = ctx->seat_staging.alloc((size_t)rows * n * 8);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->seat_staging.p, y_body, rows * n * 8, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = launch_seat(ctx->logn, (int)rows, ctx->seat_staging.as<uint64_t>(), ctx->psi.as<uint64_t>(),
                            ctx->inv_n, seat->yprime.as<Pair>(), ctx->stream);
        if (e == hipSuccess && rows == 3)
            e = launch_fold_rows(seat->yprime.as<Pair>(), (int)n, ctx->kx.w, ctx->k0.w, ctx->k1.w, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(seat->ready, ctx->stream);
    }
    if (e != hipSuccess) {
        set_err(err, err_cap, "HIP failure seating the packing key: %s", hipGetErrorString(e));
        delete seat;
        return INSPIRE_ERR_DEVICE;
    }
    *out = seat;
    return INSPIRE_OK;
}

extern "C" {

int inspire_ctx_create(const inspire_params *params, const uint8_t w_seed[32], uint32_t num_columns,
                       inspire_ctx **out, char *err, size_t err_cap) {
    if (!params || !out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    *out = nullptr;
    const uint32_t n = params->ring_dim;
    if (n != 256 && n != 512 && n != 1024 && n != 2048 && n != 4096) {
        set_err(err, err_cap,
                "ring_dim %u has no gfx950 kernel here; built ring dimensions are 256, 512, 1024, 2048 and 4096",
                n);
        return INSPIRE_ERR_INVALID;
    }
    if (params->q != kQ) {
        set_err(err, err_cap, "modulus %llu is not the single 60-bit prime 2^60 - 2^14 + 1 these kernels reduce by",
                (unsigned long long)params->q);
        return INSPIRE_ERR_INVALID;
    }
    if (params->gadget_base != (1ULL << 20) || params->packing_gadget_len != 3) {
        set_err(err, err_cap, "packing gadget (base %llu, len %u) unsupported: the kernels decompose base 2^20, 3 digits",
                (unsigned long long)params->gadget_base, params->packing_gadget_len);
        return INSPIRE_ERR_INVALID;
    }
    bool legal = num_columns != 0 && (num_columns & (num_columns - 1)) == 0 && num_columns <= n / 2 &&
                 num_columns <= 1024;
    if (!legal) {
        set_err(err, err_cap,
                "shard column count is not a legal InspiRING width: InspiRING packing width %u is illegal at ring_dim "
                "%u: the width must be a power of two in [1, %u] so that the automorphism indices form a subgroup of "
                "Z*_%u of order %u",
                num_columns, n, n / 2, 2 * n, num_columns);
        return INSPIRE_ERR_INVALID;
    }
    inspire_ctx *ctx = new (std::nothrow) inspire_ctx();
    if (!ctx) { set_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    ctx->n = (int)n;
    ctx->logn = 0;
    while ((1u << ctx->logn) < n) ctx->logn++;
    ctx->gamma = (int)num_columns;
    ctx->dropped = params->approximate_packing_gadget ? 1 : 0;
    ctx->rows = 3 - ctx->dropped;

    const uint64_t two_n = 2ULL * n;
    uint64_t psi = find_psi(two_n);
    std::vector<uint64_t> fwd = ladder(n, psi), inv = ladder(n, inv_mod(psi));
    uint64_t g = two_n / num_columns + 1;  // PackParams::partial_generator (inspiring2.rs:196-198)
    std::vector<uint64_t> pows(n);
    uint64_t v = 1;
    for (uint32_t i = 0; i < n; i++) { pows[i] = v; v = v * g % two_n; }
    std::vector<uint32_t> gpow(num_columns), gback(num_columns);
    for (uint32_t i = 0; i < num_columns; i++) {
        gpow[i] = (uint32_t)pows[i];
        gback[i] = (uint32_t)pows[(n - i) % n];
    }
    const uint64_t base = 1ULL << 20;
    uint64_t inv_b = inv_mod(base), inv_b2 = (uint64_t)((u128)inv_b * inv_b % kQ);
    ctx->inv_n = inv_mod(n);
    ctx->inv_n_gamma = inv_mod((uint64_t)((u128)n * num_columns % kQ));
    auto pair = [](uint64_t w) { return Pair{w, shoup_companion(w)}; };
    ctx->kx = pair((uint64_t)((u128)n * inv_b2 % kQ));
    ctx->k0 = pair(kQ - inv_b2);
    ctx->k1 = pair(kQ - inv_b);
    std::vector<Pair> fwd_p(n), inv_p(n);
    for (uint32_t m = 0; m < n; m++) {
        fwd_p[m] = pair(fwd[m]);
        inv_p[m] = pair(inv[m]);
    }

    hipError_t e = hipGetDevice(&ctx->device);
    if (e == hipSuccess) e = scratch_setup(ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = ctx->psi.alloc(n * 8);
    if (e == hipSuccess) e = ctx->psi_inv.alloc(n * 8);
    if (e == hipSuccess) e = ctx->gpow.alloc(num_columns * 4);
    if (e == hipSuccess) e = ctx->gpow_back.alloc(num_columns * 4);
    if (e == hipSuccess) e = ctx->twf.alloc(n * sizeof(Pair));
    if (e == hipSuccess) e = ctx->twi.alloc(n * sizeof(Pair));
    if (e == hipSuccess) e = hipMemcpy(ctx->psi.p, fwd.data(), n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->psi_inv.p, inv.data(), n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->twf.p, fwd_p.data(), n * sizeof(Pair), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->twi.p, inv_p.data(), n * sizeof(Pair), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->gpow.p, gpow.data(), num_columns * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->gpow_back.p, gback.data(), num_columns * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_err(err, err_cap, "HIP failure creating the context: %s", hipGetErrorString(e));
        delete ctx;
        return INSPIRE_ERR_DEVICE;
    }
    if (w_seed) {
        int rc = inspire_ctx_attach_crs(ctx, w_seed, err, err_cap);
        if (rc) {
            delete ctx;
            return rc;
        }
    }
    *out = ctx;
    return INSPIRE_OK;
}

int inspire_ctx_attach_crs(inspire_ctx *ctx, const uint8_t w_seed[32], char *err, size_t err_cap) {
    if (!ctx || !w_seed) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->has_crs) {
        if (memcmp(ctx->w_seed, w_seed, 32) == 0) return INSPIRE_OK;
        set_err(err, err_cap,
                "context is already bound to another CRS (w seed): its offline packing keys and every seat "
                "registered against them belong to that CRS; create a context for the new one");
        return INSPIRE_ERR_INVALID;
    }
    const size_t n = ctx->n;
    DevBuf d_seed;
    hipError_t e = ctx->wprime.alloc((size_t)ctx->rows * n * sizeof(Pair));
    if (e == hipSuccess) e = d_seed.alloc(32);
    if (e == hipSuccess) e = hipMemcpyAsync(d_seed.p, w_seed, 32, hipMemcpyHostToDevice, ctx->stream);
    // kept_mask_rows (inspiring2.rs:839-855): stream 0 for the exact gadget, gamma otherwise
    if (e == hipSuccess)
        e = launch_expand_mask(ctx->logn, ctx->rows, d_seed.as<uint8_t>(), ctx->dropped ? ctx->gamma : 0,
                               ctx->dropped, ctx->psi.as<uint64_t>(), ctx->inv_n, ctx->wprime.as<Pair>(), ctx->stream);
    if (e == hipSuccess && ctx->rows == 3)
        e = launch_fold_rows(ctx->wprime.as<Pair>(), (int)n, ctx->kx.w, ctx->k0.w, ctx->k1.w, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_err(err, err_cap, "HIP failure deriving the offline packing keys: %s", hipGetErrorString(e));
        return INSPIRE_ERR_DEVICE;
    }
    memcpy(ctx->w_seed, w_seed, 32);
    ctx->has_crs = true;
    return INSPIRE_OK;
}

void inspire_ctx_destroy(inspire_ctx *ctx) {
    if (!ctx) return;
    if (ctx->default_ws) inspire_workspace_destroy(ctx->default_ws);
    delete ctx;
}

uint32_t inspire_ctx_key_rows(const inspire_ctx *ctx) { return ctx ? (uint32_t)ctx->rows : 0; }

int inspire_ctx_set_queries_per_cu(inspire_ctx *ctx, uint32_t queries_per_cu, char *err, size_t err_cap) {
    if (!ctx || queries_per_cu < 1 || queries_per_cu > 3) {
        set_err(err, err_cap, "queries per CU must be 1, 2 or 3");
        return INSPIRE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->queries_per_cu = (int)queries_per_cu;
    return INSPIRE_OK;
}

int inspire_ctx_set_scan_splits(inspire_ctx *ctx, uint32_t splits, char *err, size_t err_cap) {
    if (!ctx) {
        set_err(err, err_cap, "null context");
        return INSPIRE_ERR_INVALID;
    }
    if (ctx->n < 2048) {
        if (splits == 1) return INSPIRE_OK;
        set_err(err, err_cap, "the DB scan below ring 2048 has one split");
        return INSPIRE_ERR_INVALID;
    }
    for (int k : kScanSplitChoices)
        if ((uint32_t)k == splits) {
            ctx->scan_splits.store(k, std::memory_order_relaxed);
            return INSPIRE_OK;
        }
    set_err(err, err_cap, "scan splits must be 8, 16 or 32 at ring %d", ctx->n);
    return INSPIRE_ERR_INVALID;
}

int inspire_shard_upload_columns(inspire_ctx *ctx, uint32_t shard_id, const uint64_t *columns, uint32_t num_columns,
                                 inspire_shard **out, char *err, size_t err_cap) {
    if (!ctx || !columns || !out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t n = ctx->n, g = ctx->gamma;
    if (num_columns != g) {
        set_err(err, err_cap,
                "respond_seeded_inspiring: shard %u column-count mismatch: got %u, expected %zu from the CRS; refusing "
                "malformed encoded state before it returns wrong plaintext",
                shard_id, num_columns, g);
        return INSPIRE_ERR_INVALID;
    }
    std::vector<uint16_t> rec(n * g);
    for (size_t j = 0; j < g; j++)
        for (size_t m = 0; m < n; m++) {
            uint64_t v = columns[j * n + m];
            if (v > 0xffff) {
                set_err(err, err_cap,
                        "shard %u column %zu coefficient %zu is %llu, not a 16-bit column value of an encoded record",
                        shard_id, j, m, (unsigned long long)v);
                return INSPIRE_ERR_INVALID;
            }
            rec[m * g + j] = (uint16_t)v;
        }
    return register_shard(ctx, shard_id, rec, out, err, err_cap);
}

int inspire_shard_upload_records(inspire_ctx *ctx, uint32_t shard_id, const uint8_t *records, size_t num_records,
                                 size_t entry_size, inspire_shard **out, char *err, size_t err_cap) {
    if (!ctx || (!records && num_records) || !out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t n = ctx->n, g = ctx->gamma;
    if ((entry_size + 1) / 2 != g) {
        set_err(err, err_cap,
                "respond_seeded_inspiring: shard %u column-count mismatch: got %zu, expected %zu from the CRS; refusing "
                "malformed encoded state before it returns wrong plaintext",
                shard_id, (entry_size + 1) / 2, g);
        return INSPIRE_ERR_INVALID;
    }
    if (num_records > n) {
        set_err(err, err_cap, "shard %u holds %zu records, more than ring_dim %zu rows", shard_id, num_records, n);
        return INSPIRE_ERR_INVALID;
    }
    std::vector<uint16_t> rec(n * g, 0);
    for (size_t m = 0; m < num_records; m++) {
        const uint8_t *e = records + m * entry_size;
        for (size_t j = 0; j < g; j++) {
            size_t o = 2 * j;
            uint16_t v = 0;
            if (o + 1 < entry_size) v = (uint16_t)(e[o] | (e[o + 1] << 8));
            else if (o < entry_size) v = e[o];
            rec[m * g + j] = v;
        }
    }
    return register_shard(ctx, shard_id, rec, out, err, err_cap);
}

int inspire_shard_precompute_packing(inspire_ctx *ctx, inspire_shard *shard, char *err, size_t err_cap) {
    if (!ctx || !shard || shard->ctx != ctx) { set_err(err, err_cap, "shard belongs to another context"); return INSPIRE_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    return build_store(ctx, shard, false, err, err_cap);
}

int inspire_shard_has_store(const inspire_shard *shard) { return shard && shard->has_store ? 1 : 0; }

int inspire_shard_write_value(inspire_shard *shard, uint32_t record, uint32_t column, uint16_t value, char *err,
                              size_t err_cap) {
    if (!shard) { set_err(err, err_cap, "null shard"); return INSPIRE_ERR_INVALID; }
    inspire_ctx *ctx = shard->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (record >= (uint32_t)ctx->n || column >= (uint32_t)ctx->gamma) {
        set_err(err, err_cap, "record %u column %u outside the %d x %d shard", record, column, ctx->n, ctx->gamma);
        return INSPIRE_ERR_INVALID;
    }
    size_t off = ((size_t)record * ctx->gamma + column) * 2;
    HIP_TRY(hipMemcpy(static_cast<char *>(shard->records.p) + off, &value, 2, hipMemcpyHostToDevice));
    if (shard->has_store && shard->lazy_store) {
        // a store the shard never asked for: forget it, so the next respond packs the edited
        // columns as the reference's store-less path does (its buffers are reused by the rebuild)
        shard->has_store = false;
        shard->lazy_store = false;
        if (shard->id < ctx->shards.size() && ctx->owners[shard->id] == shard) {
            ctx->shards[shard->id].store = nullptr;
            ctx->table_dirty = true;
            int rc = sync_table(ctx, err, err_cap);
            if (rc) return rc;
        }
    }
    return INSPIRE_OK;
}

size_t inspire_shard_resident_bytes(const inspire_shard *shard) {
    if (!shard) return 0;
    return shard->records.bytes + (shard->has_store ? shard->store.bytes + shard->tag_key.bytes + shard->tags.bytes : 0);
}

void inspire_shard_destroy(inspire_shard *shard) {
    if (!shard) return;
    inspire_ctx *ctx = shard->ctx;
    // Two phases. First the host forgets the shard (host-buffer calls naming it are refused)
    // and every launch enqueued so far finishes with the shard as it was: those launches were
    // enqueued under the lock with a fence event behind them, and store builds and host-buffer
    // responds wait for their own stream before they release it.
    std::vector<hipEvent_t> pending;
    bool owned = false;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        owned = shard->id < ctx->owners.size() && ctx->owners[shard->id] == shard;
        if (owned) ctx->owners[shard->id] = nullptr;
        pending = ctx->fence.snapshot();
    }
    (void)wait_for(pending);
    // Then the device table drops it, so launches from here on get status -1 for it; the ones
    // enqueued between the two phases may have read it either way and are waited for too.
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        // (unless a new shard was uploaded at the id meanwhile)
        if (owned && ctx->shards[shard->id].records == shard->records.as<uint16_t>()) {
            ctx->shards[shard->id] = ShardDev{nullptr, nullptr, nullptr, nullptr};
            ctx->table_dirty = true;
            char e[160];
            (void)sync_table(ctx, e, sizeof e);
        }
        pending = ctx->fence.snapshot();
    }
    (void)wait_for(pending);
    delete shard;
}

int inspire_seat_create(inspire_ctx *ctx, const uint64_t *y_body, uint32_t rows, uint32_t num_to_pack,
                        inspire_seat **out, char *err, size_t err_cap) {
    inspire_seat *seat = nullptr;
    int rc = seat_create_enqueue(ctx, y_body, rows, num_to_pack, &seat, err, err_cap);
    if (rc) return rc;
    hipError_t e = hipEventSynchronize(seat->ready);
    if (e != hipSuccess) {
        set_err(err, err_cap, "HIP failure seating the packing key: %s", hipGetErrorString(e));
        inspire_seat_destroy(seat);
        return INSPIRE_ERR_DEVICE;
    }
    seat->settled = true;
    *out = seat;
    return INSPIRE_OK;
}

size_t inspire_seat_resident_bytes(const inspire_seat *seat) { return seat ? seat->yprime.bytes : 0; }

// The launches enqueued so far on ctx (those that may read an object about to be freed).
static std::vector<hipEvent_t> launches_so_far(inspire_ctx *ctx) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return ctx->fence.snapshot();
}

void inspire_seat_destroy(inspire_seat *seat) {
    if (!seat) return;
    inspire_ctx *ctx = seat->ctx;
    (void)wait_for(launches_so_far(ctx));
    {
        // no launch enqueued so far reads the seat any more: its memory goes back to the pool
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->seat_pool.give(seat->yprime.p);
        seat->yprime.p = nullptr;
    }
    delete seat;
}

int inspire_workspace_create(inspire_ctx *ctx, size_t max_count, inspire_workspace **out, char *err, size_t err_cap) {
    if (!ctx || !out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    inspire_workspace *ws = new (std::nothrow) inspire_workspace();
    if (!ws) { set_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    ws->ctx = ctx;
    int rc = ws_reserve(ws, max_count, err, err_cap);
    if (rc) { delete ws; return rc; }
    *out = ws;
    return INSPIRE_OK;
}

void inspire_workspace_destroy(inspire_workspace *ws) {
    if (!ws) return;
    (void)wait_for(launches_so_far(ws->ctx));
    delete ws;
}

static int respond_device(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                          const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b, size_t count,
                          uint64_t switch_modulus, uint64_t *d_a_out, uint64_t *d_b_prefix_out, void *stream,
                          float *timing, char *err, size_t err_cap, uint8_t *d_wire = nullptr,
                          size_t wire_stride = 0, int queries_per_cu = 0, int scan_splits = 0) {
    if (!ctx || !ws || ws->ctx != ctx) { set_err(err, err_cap, "workspace belongs to another context"); return INSPIRE_ERR_INVALID; }
    int rc = check_batch(ctx, seat, nullptr, count, err, err_cap);
    if (rc) return rc;
    if (count > ws->max_count) {
        set_err(err, err_cap, "batch of %zu queries exceeds the workspace's %zu", count, ws->max_count);
        return INSPIRE_ERR_INVALID;
    }
    if (switch_modulus >= kSwitchModulusLimit) {
        set_err(err, err_cap, "switch modulus %llu is not below 2^46 (the 36- and 45-bit targets are implemented)",
                (unsigned long long)switch_modulus);
        return INSPIRE_ERR_INVALID;
    }
    BatchDev b;
    b.count = (int)count;
    b.shard_ids = d_shard_ids;
    b.seeds = d_seeds;
    b.yprimes = nullptr;
    b.a_in = nullptr;
    b.b = d_b;
    b.yprime = seat->yprime.as<Pair>();
    b.scan_part = ws->scan_part.as<uint64_t>();
    b.ks = scan_splits ? scan_splits : ctx->ks();  // a trial tile (inspire_ctx_autotune_scan)
    b.a_out = d_a_out;
    b.b_out = d_b_prefix_out;
    b.status = ws->status.as<int32_t>();
    b.switch_modulus = switch_modulus;
    ws_group(ws, b);
    if (d_wire) {
        const WireLayout wl = wire_layout((uint32_t)ctx->n, (uint32_t)ctx->gamma, switch_modulus ? switch_modulus : kQ);
        if (wire_stride % 4 || wire_stride < ((wl.total + 3) & ~3u) || reinterpret_cast<uintptr_t>(d_wire) % 4) {
            set_err(err, err_cap,
                    "response buffer must be 4-byte aligned with a stride that is a multiple of 4 and at least %u "
                    "(the %u-byte response rounded up)",
                    (wl.total + 3) & ~3u, wl.total);
            return INSPIRE_ERR_INVALID;
        }
        b.wire_out = d_wire;
        b.wire_stride = wire_stride;
    } else if (!d_a_out || !d_b_prefix_out) {
        set_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the shard table is read and the kernels are enqueued under the context's lock, with a
    // fence event behind them, so destroying a shard, seat or workspace waits for this launch
    std::unique_lock<std::mutex> lk(ctx->mu);
    rc = sync_table(ctx, err, err_cap);
    if (rc) return rc;
    CtxDev dev = ctx->dev();
    if (queries_per_cu) dev.queries_per_cu = queries_per_cu;  // a trial shape (autotune)
    if (!timing) {
        if (hipEvent_t *tr = ws->trace.next()) {
            // a traced launch: events around the grouping, scan and packing kernels, read later,
            // no wait
            hipError_t e = hipEventRecord(tr[0], st);
            if (e == hipSuccess) e = launch_group_only(dev, b, st);
            if (e == hipSuccess) e = hipEventRecord(tr[1], st);
            if (e == hipSuccess) e = launch_scan_only(dev, b, st);
            if (e == hipSuccess) e = hipEventRecord(tr[2], st);
            if (e == hipSuccess) e = launch_chain_only(dev, b, st);
            if (e == hipSuccess) e = hipEventRecord(tr[3], st);
            HIP_TRY(e);
            ws->trace.done();
        } else {
            HIP_TRY(launch_respond(dev, b, st));
        }
        HIP_TRY(ctx->fence.note(st));
        return INSPIRE_OK;
    }
    // per-kernel durations from events on the launch stream
    hipEvent_t ev[3];
    for (int i = 0; i < 3; i++) HIP_TRY(hipEventCreate(&ev[i]));
    hipError_t e = hipEventRecord(ev[0], st);
    if (e == hipSuccess) e = launch_group_only(dev, b, st);
    if (e == hipSuccess) e = launch_scan_only(dev, b, st);
    if (e == hipSuccess) e = hipEventRecord(ev[1], st);
    if (e == hipSuccess) e = launch_chain_only(dev, b, st);
    if (e == hipSuccess) e = hipEventRecord(ev[2], st);
    if (e == hipSuccess) e = ctx->fence.note(st);
    lk.unlock();
    if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&timing[0], ev[0], ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&timing[1], ev[1], ev[2]);
    for (int i = 0; i < 3; i++) (void)hipEventDestroy(ev[i]);
    HIP_TRY(e);
    return INSPIRE_OK;
}

int inspire_respond_batch_device(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                 const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                 size_t count, uint64_t switch_modulus, uint64_t *d_a_out, uint64_t *d_b_prefix_out,
                                 void *stream, char *err, size_t err_cap) {
    return respond_device(ctx, seat, ws, d_shard_ids, d_seeds, d_b, count, switch_modulus, d_a_out, d_b_prefix_out,
                          stream, nullptr, err, err_cap);
}

int inspire_respond_batch_device_noalloc(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *d_shard_ids,
                                         const uint8_t *d_seeds, const uint64_t *d_b, size_t count,
                                         uint64_t switch_modulus, uint64_t *d_a_out, uint64_t *d_b_prefix_out,
                                         int32_t *d_status, void *stream, char *err, size_t err_cap) {
    if (!ctx || (count && (!d_shard_ids || !d_seeds || !d_b || !d_a_out || !d_b_prefix_out || !d_status))) {
        set_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    int rc = check_batch(ctx, seat, nullptr, count, err, err_cap);
    if (rc) return rc;
    if (switch_modulus >= kSwitchModulusLimit) {
        set_err(err, err_cap, "switch modulus %llu is not below 2^46 (the 36- and 45-bit targets are implemented)",
                (unsigned long long)switch_modulus);
        return INSPIRE_ERR_INVALID;
    }
    const int ks = ctx->ks();  // read once: the limit and the launch agree if a tuning runs meanwhile
    const size_t per_query = (size_t)ks * ctx->gamma * 4 * sizeof(uint64_t);
    if (count > INSPIRE_NOALLOC_SCRATCH_BYTES / per_query) {
        set_err(err, err_cap,
                "batch of %zu queries needs %zu bytes of scan scratch; the allocation-free entry point holds %u "
                "bytes per GPU, at most %zu queries at this ring and width",
                count, count * per_query, (unsigned)INSPIRE_NOALLOC_SCRATCH_BYTES,
                (size_t)INSPIRE_NOALLOC_SCRATCH_BYTES / per_query);
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0) return INSPIRE_OK;
    void *scr = scratch_of(ctx->device);
    if (!scr) {
        set_err(err, err_cap, "no allocation-free scratch on device %d", ctx->device);
        return INSPIRE_ERR_DEVICE;
    }
    CtxDev dev;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        // the device shard table is refreshed in place when it fits (a copy, no allocation); a
        // table that has outgrown its buffer needs one allocating call first
        if (ctx->table_dirty && (!ctx->shard_table.p || ctx->shard_table.bytes < ctx->shards.size() * sizeof(ShardDev))) {
            set_err(err, err_cap,
                    "the shard table outgrew its device buffer since the last launch; make one allocating call "
                    "(a respond or inspire_shard_precompute_packing) before the allocation-free one");
            return INSPIRE_ERR_INVALID;
        }
        rc = sync_table(ctx, err, err_cap);
        if (rc) return rc;
        dev = ctx->dev();
    }
    BatchDev b;
    b.count = (int)count;
    b.shard_ids = d_shard_ids;
    b.seeds = d_seeds;
    b.yprimes = nullptr;
    b.a_in = nullptr;
    b.b = d_b;
    b.yprime = seat->yprime.as<Pair>();
    b.scan_part = static_cast<uint64_t *>(scr);
    b.ks = ks;
    b.a_out = d_a_out;
    b.b_out = d_b_prefix_out;
    b.status = d_status;
    b.switch_modulus = switch_modulus;
    HIP_TRY(launch_respond(dev, b, static_cast<hipStream_t>(stream)));
    return INSPIRE_OK;
}

int inspire_workspace_trace(inspire_workspace *ws, size_t launches, char *err, size_t err_cap) {
    if (!ws) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    if (launches > (1u << 20)) {
        set_err(err, err_cap, "a trace holds at most %u launches, asked for %zu", 1u << 20, launches);
        return INSPIRE_ERR_INVALID;
    }
    HIP_TRY(ws->trace.arm(launches, 3));
    return INSPIRE_OK;
}

int inspire_workspace_kernel_times(inspire_workspace *ws, float *group_ms, float *scan_ms, float *chain_ms,
                                   size_t cap, size_t *count, char *err, size_t err_cap) {
    if (!ws || !count || (cap && (!group_ms || !scan_ms || !chain_ms))) {
        set_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    float *const out[3] = {group_ms, scan_ms, chain_ms};
    HIP_TRY(ws->trace.read(out, 3, cap, count));
    return INSPIRE_OK;
}

int inspire_respond_batch_device_timed(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                       const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                       size_t count, uint64_t switch_modulus, uint64_t *d_a_out,
                                       uint64_t *d_b_prefix_out, void *stream, float kernel_ms[2], char *err,
                                       size_t err_cap) {
    if (!kernel_ms) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    return respond_device(ctx, seat, ws, d_shard_ids, d_seeds, d_b, count, switch_modulus, d_a_out, d_b_prefix_out,
                          stream, kernel_ms, err, err_cap);
}

static int respond_host(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids, const uint8_t *seeds,
                        const uint64_t *a_in, const uint64_t *b, size_t count, uint64_t switch_modulus,
                        uint64_t *a_out, uint64_t *b_prefix_out, char *err, size_t err_cap,
                        uint8_t *wire_out = nullptr, size_t wire_stride = 0);

size_t inspire_response_wire_bytes(const inspire_ctx *ctx, uint64_t switch_modulus) {
    if (!ctx) return 0;
    return wire_layout((uint32_t)ctx->n, (uint32_t)ctx->gamma, switch_modulus ? switch_modulus : kQ).total;
}

int inspire_respond_batch_device_wire(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                      const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                      size_t count, uint64_t switch_modulus, uint8_t *d_wire, size_t wire_stride,
                                      void *stream, char *err, size_t err_cap) {
    if (count && !d_wire) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    return respond_device(ctx, seat, ws, d_shard_ids, d_seeds, d_b, count, switch_modulus, nullptr, nullptr, stream,
                          nullptr, err, err_cap, d_wire, wire_stride);
}

int inspire_respond_batch_wire(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids,
                               const uint8_t *seeds, const uint64_t *b, size_t count, uint64_t switch_modulus,
                               uint8_t *wire_out, size_t wire_stride, char *err, size_t err_cap) {
    if (count && (!seeds || !wire_out)) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    return respond_host(ctx, seat, shard_ids, seeds, nullptr, b, count, switch_modulus, nullptr, nullptr, err, err_cap,
                        wire_out, wire_stride);
}

struct inspire_graph {
    inspire_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~inspire_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

int inspire_respond_graph_create(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                 const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                 size_t count, uint64_t switch_modulus, uint64_t *d_a_out, uint64_t *d_b_prefix_out,
                                 inspire_graph **out, char *err, size_t err_cap) {
    if (!out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    if (!ctx || !ws || ws->ctx != ctx) { set_err(err, err_cap, "workspace belongs to another context"); return INSPIRE_ERR_INVALID; }
    int rc = check_batch(ctx, seat, nullptr, count, err, err_cap);
    if (rc) return rc;
    if (count == 0 || count > ws->max_count) {
        set_err(err, err_cap, "batch of %zu queries does not fit the workspace's %zu", count, ws->max_count);
        return INSPIRE_ERR_INVALID;
    }
    if (switch_modulus >= kSwitchModulusLimit) {
        set_err(err, err_cap, "switch modulus %llu is not below 2^46 (the 36- and 45-bit targets are implemented)",
                (unsigned long long)switch_modulus);
        return INSPIRE_ERR_INVALID;
    }
    CtxDev dev;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rc = sync_table(ctx, err, err_cap);
        if (rc) return rc;
        dev = ctx->dev();
    }
    BatchDev b;
    b.count = (int)count;
    b.shard_ids = d_shard_ids;
    b.seeds = d_seeds;
    b.yprimes = nullptr;
    b.a_in = nullptr;
    b.b = d_b;
    b.yprime = seat->yprime.as<Pair>();
    b.scan_part = ws->scan_part.as<uint64_t>();
    b.ks = ctx->ks();
    b.a_out = d_a_out;
    b.b_out = d_b_prefix_out;
    b.status = ws->status.as<int32_t>();
    b.switch_modulus = switch_modulus;
    ws_group(ws, b);
    inspire_graph *g = new (std::nothrow) inspire_graph();
    if (!g) { set_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    g->ctx = ctx;
    // the two kernels of the pipeline, captured on a private stream (no memset or copy nodes)
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = launch_respond(dev, b, cs);
        e = hipStreamEndCapture(cs, &g->graph);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        set_err(err, err_cap, "HIP failure capturing the respond graph: %s", hipGetErrorString(e));
        delete g;
        return INSPIRE_ERR_DEVICE;
    }
    *out = g;
    return INSPIRE_OK;
}

int inspire_graph_launch(inspire_graph *g, void *stream, char *err, size_t err_cap) {
    if (!g) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(g->ctx->mu);
    HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    HIP_TRY(g->ctx->fence.note(static_cast<hipStream_t>(stream)));
    return INSPIRE_OK;
}

void inspire_graph_destroy(inspire_graph *g) { delete g; }

// Synthetic queries of the real shape for the autotuners: ids over the shards with a store
// (round robin, or uniform draws), distinct seeds, b uniform-looking below q; device copies.
struct TuneBatch {
    DevBuf ids, seeds, b, a, bp;
};
static int tune_batch(inspire_ctx *ctx, size_t count, bool uniform, TuneBatch &tb, char *err, size_t err_cap) {
    std::vector<uint32_t> ids;
    uint64_t x = 0x9e3779b97f4a7c15ULL;
    auto next = [&x]() { x ^= x << 13; x ^= x >> 7; x ^= x << 17; return x; };
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        std::vector<uint32_t> stored;
        for (size_t i = 0; i < ctx->owners.size(); i++)
            if (ctx->owners[i] && ctx->owners[i]->has_store) stored.push_back((uint32_t)i);
        if (stored.empty()) {
            set_err(err, err_cap, "autotune needs a shard with its packing store built");
            return INSPIRE_ERR_INVALID;
        }
        for (size_t k = 0; k < count; k++) ids.push_back(stored[uniform ? next() % stored.size() : k % stored.size()]);
    }
    const size_t n = ctx->n, g = ctx->gamma;
    std::vector<uint8_t> seeds(count * 32);
    std::vector<uint64_t> b(count * n);
    for (auto &v : seeds) v = (uint8_t)next();
    for (auto &v : b) v = next() % kQ;
    HIP_TRY(tb.ids.alloc(count * 4));
    HIP_TRY(tb.seeds.alloc(count * 32));
    HIP_TRY(tb.b.alloc(count * n * 8));
    HIP_TRY(tb.a.alloc(count * n * 8));
    HIP_TRY(tb.bp.alloc(count * g * 8));
    HIP_TRY(hipMemcpy(tb.ids.p, ids.data(), count * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tb.seeds.p, seeds.data(), count * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tb.b.p, b.data(), count * n * 8, hipMemcpyHostToDevice));
    return INSPIRE_OK;
}

static int tune_args(inspire_ctx *ctx, inspire_workspace *ws, size_t count, const void *out_a, const void *out_b,
                     char *err, size_t err_cap) {
    if (!ctx || !ws || ws->ctx != ctx || !out_a || !out_b) {
        set_err(err, err_cap, "null argument or workspace of another context");
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0 || count > ws->max_count) {
        set_err(err, err_cap, "batch of %zu queries does not fit the workspace's %zu", count, ws->max_count);
        return INSPIRE_ERR_INVALID;
    }
    return INSPIRE_OK;
}

int inspire_ctx_autotune(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws, size_t count,
                         float kernel_ms[3], uint32_t *chosen, char *err, size_t err_cap) {
    int rc = tune_args(ctx, ws, count, kernel_ms, chosen, err, err_cap);
    if (rc) return rc;
    TuneBatch tb;
    rc = tune_batch(ctx, count, false, tb, err, err_cap);
    if (rc) return rc;
    // each shape is tried on its own launches; the context's shape changes only at the end,
    // so launches from other threads meanwhile keep the shape they had
    float best = 0;
    uint32_t pick = 1;
    for (int qpc = 1; qpc <= 3; qpc++) {
        float t[2] = {0, 0}, sum = 0;
        for (int rep = 0; rep < 3; rep++) {  // one warm-up, two timed
            rc = respond_device(ctx, seat, ws, tb.ids.as<uint32_t>(), tb.seeds.as<uint8_t>(), tb.b.as<uint64_t>(),
                                count, 0, tb.a.as<uint64_t>(), tb.bp.as<uint64_t>(), nullptr, t, err, err_cap,
                                nullptr, 0, qpc);
            if (rc) return rc;
            if (rep) sum += t[0] + t[1];
        }
        kernel_ms[qpc - 1] = sum / 2;
        if (qpc == 1 || sum / 2 < best) { best = sum / 2; pick = (uint32_t)qpc; }
    }
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->queries_per_cu = (int)pick;
    }
    *chosen = pick;
    return INSPIRE_OK;
}

int inspire_ctx_autotune_scan(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws, size_t count,
                              float scan_ms[3], uint32_t *chosen, char *err, size_t err_cap) {
    int rc = tune_args(ctx, ws, count, scan_ms, chosen, err, err_cap);
    if (rc) return rc;
    if (ctx->n < 2048) {  // one split: nothing to choose
        for (int k = 0; k < 3; k++) scan_ms[k] = 0;
        *chosen = 1;
        return INSPIRE_OK;
    }
    // shard ids drawn uniformly, as a launch's queries arrive: the tasks of the grouped scan
    // (shards named by one, two or more queries) have the mix real launches have
    TuneBatch tb;
    rc = tune_batch(ctx, count, true, tb, err, err_cap);
    if (rc) return rc;
    float best = 0;
    int pick = kScanSplitChoices[0];
    for (int k = 0; k < 3; k++) {
        const int ks = kScanSplitChoices[k];
        float t[2] = {0, 0}, sum = 0;
        for (int rep = 0; rep < 3; rep++) {  // one warm-up, two timed
            rc = respond_device(ctx, seat, ws, tb.ids.as<uint32_t>(), tb.seeds.as<uint8_t>(), tb.b.as<uint64_t>(),
                                count, 0, tb.a.as<uint64_t>(), tb.bp.as<uint64_t>(), nullptr, t, err, err_cap,
                                nullptr, 0, 0, ks);
            if (rc) return rc;
            if (rep) sum += t[0];  // grouping and scan kernels
        }
        scan_ms[k] = sum / 2;
        if (k == 0 || sum / 2 < best) { best = sum / 2; pick = ks; }
    }
    ctx->scan_splits.store(pick, std::memory_order_relaxed);
    *chosen = (uint32_t)pick;
    return INSPIRE_OK;
}

int inspire_workspace_status(inspire_workspace *ws, size_t count, int32_t *status, char *err, size_t err_cap) {
    if (!ws || !status || count > ws->max_count) { set_err(err, err_cap, "bad status request"); return INSPIRE_ERR_INVALID; }
    HIP_TRY(hipMemcpy(status, ws->status.p, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    return INSPIRE_OK;
}

static int respond_host(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids, const uint8_t *seeds,
                        const uint64_t *a_in, const uint64_t *b, size_t count, uint64_t switch_modulus,
                        uint64_t *a_out, uint64_t *b_prefix_out, char *err, size_t err_cap,
                        uint8_t *wire_out, size_t wire_stride) {
    if (!ctx || (count && (!shard_ids || (!seeds && !a_in) || !b || (!wire_out && (!a_out || !b_prefix_out))))) {
        set_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    // the seat is checked even for an empty batch, as respond_seeded_batch resolves its key
    // before mapping over the queries (respond.rs:1189-1202)
    int rc = check_batch(ctx, seat, shard_ids, count, err, err_cap);
    if (rc) return rc;
    if (count == 0) return INSPIRE_OK;
    const size_t n = ctx->n, g = ctx->gamma;
    for (size_t k = 0; k < count; k++)
        for (size_t j = 0; j < n; j++)
            if (b[k * n + j] >= kQ || (a_in && a_in[k * n + j] >= kQ)) {
                set_err(err, err_cap,
                        "Poly decode refused: limb 0 coefficient %zu is %llu, not canonical for modulus %llu", j,
                        (unsigned long long)(b[k * n + j] >= kQ ? b[k * n + j] : a_in[k * n + j]),
                        (unsigned long long)kQ);
                return INSPIRE_ERR_INVALID;
            }
    if (switch_modulus >= kSwitchModulusLimit) {
        set_err(err, err_cap, "switch modulus %llu is not below 2^46 (the 36- and 45-bit targets are implemented)",
                (unsigned long long)switch_modulus);
        return INSPIRE_ERR_INVALID;
    }
    // a shard decoded without its store packs through one built now: the bytes equal the
    // per-query path's (shard_store.rs:1-16), so building lazily changes nothing observable
    for (size_t k = 0; k < count; k++) {
        inspire_shard *sh = ctx->owners[shard_ids[k]];
        if (!sh->has_store) {
            rc = build_store(ctx, sh, true, err, err_cap);
            if (rc) return rc;
        }
    }
    rc = sync_table(ctx, err, err_cap);
    if (rc) return rc;
    if (!ctx->default_ws) {
        rc = inspire_workspace_create(ctx, count, &ctx->default_ws, err, err_cap);
        if (rc) return rc;
    }
    inspire_workspace *ws = ctx->default_ws;
    if (count > ws->max_count || !ws->ids.p) {
        rc = ws_reserve(ws, count, err, err_cap);
        if (rc) return rc;
        HIP_TRY(ws->ids.alloc(ws->max_count * 4));
        HIP_TRY(ws->seeds.alloc(ws->max_count * 32));
        HIP_TRY(ws->a_in.alloc(ws->max_count * n * 8));
        HIP_TRY(ws->b.alloc(ws->max_count * n * 8));
        HIP_TRY(ws->a_out.alloc(ws->max_count * n * 8));
        HIP_TRY(ws->b_out.alloc(ws->max_count * g * 8));
    }
    HIP_TRY(hipMemcpy(ws->ids.p, shard_ids, count * 4, hipMemcpyHostToDevice));
    if (seeds) HIP_TRY(hipMemcpy(ws->seeds.p, seeds, count * 32, hipMemcpyHostToDevice));
    if (a_in) HIP_TRY(hipMemcpy(ws->a_in.p, a_in, count * n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws->b.p, b, count * n * 8, hipMemcpyHostToDevice));
    BatchDev bd;
    bd.count = (int)count;
    bd.shard_ids = ws->ids.as<uint32_t>();
    bd.seeds = ws->seeds.as<uint8_t>();
    bd.a_in = a_in ? ws->a_in.as<uint64_t>() : nullptr;
    bd.yprimes = nullptr;
    bd.b = ws->b.as<uint64_t>();
    bd.yprime = seat->yprime.as<Pair>();
    bd.scan_part = ws->scan_part.as<uint64_t>();
    bd.ks = ctx->ks();
    bd.a_out = ws->a_out.as<uint64_t>();
    bd.b_out = ws->b_out.as<uint64_t>();
    bd.status = ws->status.as<int32_t>();
    bd.switch_modulus = switch_modulus;
    ws_group(ws, bd);
    const WireLayout wl = wire_layout((uint32_t)n, (uint32_t)g, switch_modulus ? switch_modulus : kQ);
    const size_t dstride = (wl.total + 3) & ~(size_t)3;
    if (wire_out) {
        if (wire_stride < wl.total) {
            set_err(err, err_cap, "response stride %zu is below the %u-byte response", wire_stride, wl.total);
            return INSPIRE_ERR_INVALID;
        }
        if (ws->wire_cap < count * dstride) {
            HIP_TRY(ws->wire.alloc(count * dstride));
            ws->wire_cap = count * dstride;
        }
        bd.wire_out = ws->wire.as<uint8_t>();
        bd.wire_stride = dstride;
    }
    HIP_TRY(launch_respond(ctx->dev(), bd, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> st(count);
    HIP_TRY(hipMemcpy(st.data(), ws->status.p, count * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < count; k++)
        if (st[k] < 0) {
            set_err(err, err_cap, "shard %u has no packing store on the device", shard_ids[k]);
            return INSPIRE_ERR_STORE;
        } else if (st[k] != 0) {
            set_err(err, err_cap,
                    "shard %u: packing store refused: column %d is not the column the store was built from. The "
                    "shard changed after the build; rebuild the store before serving it",
                    shard_ids[k], st[k] - 1);
            return INSPIRE_ERR_STORE;
        }
    if (a_out) HIP_TRY(hipMemcpy(a_out, ws->a_out.p, count * n * 8, hipMemcpyDeviceToHost));
    if (b_prefix_out) HIP_TRY(hipMemcpy(b_prefix_out, ws->b_out.p, count * g * 8, hipMemcpyDeviceToHost));
    if (wire_out)
        HIP_TRY(hipMemcpy2D(wire_out, wire_stride, ws->wire.p, dstride, wl.total, count, hipMemcpyDeviceToHost));
    return INSPIRE_OK;
}

int inspire_respond_batch(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids, const uint8_t *seeds,
                          const uint64_t *b, size_t count, uint64_t switch_modulus, uint64_t *a_out,
                          uint64_t *b_prefix_out, char *err, size_t err_cap) {
    if (count && !seeds) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    return respond_host(ctx, seat, shard_ids, seeds, nullptr, b, count, switch_modulus, a_out, b_prefix_out, err,
                        err_cap);
}

int inspire_respond_batch_expanded(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids,
                                   const uint64_t *a, const uint64_t *b, size_t count, uint64_t switch_modulus,
                                   uint64_t *a_out, uint64_t *b_prefix_out, char *err, size_t err_cap) {
    if (count && !a) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    return respond_host(ctx, seat, shard_ids, nullptr, a, b, count, switch_modulus, a_out, b_prefix_out, err,
                        err_cap);
}

}  // extern "C"
