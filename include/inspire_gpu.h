/*
 * inspire_gpu.h - C ABI of the MI355X (gfx950) InsPIRe respond path.
 *
 * Plain pointers and sizes only. Every entry point returns INSPIRE_OK (0) or an
 * error code and, on error, writes a NUL-terminated message into err[0..err_cap)
 * (the reference's PirError(String), src/pir/error.rs:6). Refusals happen on the
 * host before any device work, with the reference's message text where the same
 * condition exists there.
 *
 * Polynomials are arrays of ring_dim u64 coefficients in [0, q), coefficient form,
 * limb-major as the reference's Poly::coeffs() (src/math/poly.rs:51-58).
 *
 * Object lifetimes: a context owns its shards' device table; shards and seats must
 * be destroyed before their context. Calls on one context are serialised by an
 * internal lock, except inspire_respond_batch_device, which only reads the context
 * and may run concurrently on distinct workspaces and streams.
 */
#ifndef INSPIRE_GPU_H
#define INSPIRE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INSPIRE_OK 0
#define INSPIRE_ERR_INVALID 1 /* refused input: shape, width, modulus, unknown shard */
#define INSPIRE_ERR_DEVICE 2  /* HIP runtime failure */
#define INSPIRE_ERR_STORE 3   /* a shard's columns no longer match its packing store */

#define INSPIRE_Q 1152921504606830593ULL      /* 2^60 - 2^14 + 1, src/math/mod_q.rs:7 */
#define INSPIRE_Q36 68718428161ULL            /* MOD_SWITCH_TARGET_36BIT, src/pir/mod_switch.rs:47 */
/* the most queries one call, workspace or launch takes (2^22): beyond it a call is refused */
#define INSPIRE_MAX_BATCH 4194304u

typedef struct inspire_ctx inspire_ctx;
typedef struct inspire_shard inspire_shard;
typedef struct inspire_seat inspire_seat;
typedef struct inspire_workspace inspire_workspace;

/* The fields of InspireParams the respond path reads (src/params.rs:120-190). */
typedef struct {
    uint32_t ring_dim;                    /* 256, 2048 or 4096 (InspireParams::d4096) */
    uint64_t q;                           /* must be INSPIRE_Q */
    uint64_t gadget_base;                 /* must be 2^20 */
    uint32_t packing_gadget_len;          /* must be 3 */
    uint32_t approximate_packing_gadget;  /* 1: the key has no row for the lowest digit */
} inspire_params;

/* PackParams::try_new for one shard width (src/inspiring/inspiring2.rs:171-286): the ring
 * tables every store and launch reads; num_columns is gamma. With w_seed (the CRS's seed of
 * the packing mask), also OfflinePackingKeys::generate, i.e. ServerInspiringCache::new
 * (src/pir/respond.rs:926-941, inspiring2.rs:537-598). With w_seed NULL the context holds no
 * CRS: it takes shards and builds their packing stores, as ShardData::precompute_packing
 * (src/pir/setup.rs:335-340) needs only PackParams and the NttContext, and every respond
 * entry point and inspire_server_create refuse it until inspire_ctx_attach_crs. */
int inspire_ctx_create(const inspire_params *params, const uint8_t w_seed[32], uint32_t num_columns,
                       inspire_ctx **out, char *err, size_t err_cap);
/* ServerInspiringCache::new(&crs, &encoded) over a context created without a CRS: derives the
 * offline packing keys from w_seed. Binding the seed the context already holds is a no-op;
 * another seed is refused (the context's seats belong to its CRS). */
int inspire_ctx_attach_crs(inspire_ctx *ctx, const uint8_t w_seed[32], char *err, size_t err_cap);
void inspire_ctx_destroy(inspire_ctx *ctx);
/* key rows the seats must carry: packing_gadget_len minus the dropped digits
 * (PackParams::key_rows, src/inspiring/inspiring2.rs:289-291) */
uint32_t inspire_ctx_key_rows(const inspire_ctx *ctx);
/* Launch shaping of the packing kernel (no reference counterpart; results are
 * unchanged): 1 (default) gives every query a CU of its own, 2 lets two queries share
 * a CU, which raises throughput when far more queries are in flight than CUs, and 3
 * runs a kernel built for three per CU (fewer registers). inspire_ctx_autotune picks
 * one by timing. */
int inspire_ctx_set_queries_per_cu(inspire_ctx *ctx, uint32_t queries_per_cu, char *err, size_t err_cap);
/* Row blocks per shard of the DB scan (no reference counterpart; results are unchanged):
 * 8, 16 (default) or 32 at ring 2048 and up, 1 below; inspire_ctx_autotune_scan picks one
 * by timing. The next launch of every workspace and serving slot of the context uses it. */
int inspire_ctx_set_scan_splits(inspire_ctx *ctx, uint32_t splits, char *err, size_t err_cap);

/* Upload one shard of the EncodedDatabase (ShardData, src/pir/setup.rs:280-289).
 * columns: num_columns x ring_dim column polynomials (ShardData::polynomials), each
 * coefficient a 16-bit column value (src/pir/encode_db.rs:76-159). The shard id is
 * its dense position, as respond's require_dense_shard demands
 * (src/pir/respond.rs:442-462). */
int inspire_shard_upload_columns(inspire_ctx *ctx, uint32_t shard_id, const uint64_t *columns,
                                 uint32_t num_columns, inspire_shard **out, char *err, size_t err_cap);
/* The same from raw records: record m occupies records[m*entry_size ..], column j of
 * record m is entry[2j] | entry[2j+1] << 8 (extract_column_value,
 * src/pir/encode_db.rs:146-159). num_records <= ring_dim; missing rows are zero. */
int inspire_shard_upload_records(inspire_ctx *ctx, uint32_t shard_id, const uint8_t *records,
                                 size_t num_records, size_t entry_size, inspire_shard **out, char *err,
                                 size_t err_cap);
/* ShardData::precompute_packing (src/pir/setup.rs:335-340): build the step-1 store
 * on the device, with keyed column tags under a fresh OS-entropy key. */
int inspire_shard_precompute_packing(inspire_ctx *ctx, inspire_shard *shard, char *err, size_t err_cap);
int inspire_shard_has_store(const inspire_shard *shard);
/* Overwrite one column value in place without rebuilding the store (a shard edited
 * after its store was built; tests/shard_packing_store.rs:328-368). After
 * inspire_shard_precompute_packing, the next respond against this shard is refused with
 * INSPIRE_ERR_STORE. A shard that never had precompute_packing (its store, if any, was built by
 * a host-buffer respond on the way) is packed from its edited columns on the next respond, as
 * the reference packs a store-less shard; device-buffer launches see no store for it (status -1)
 * until then. */
int inspire_shard_write_value(inspire_shard *shard, uint32_t record, uint32_t column, uint16_t value, char *err,
                              size_t err_cap);
/* HBM bytes held by the shard: records + store + tag key + tags. */
size_t inspire_shard_resident_bytes(const inspire_shard *shard);
void inspire_shard_destroy(inspire_shard *shard);

/* PackingKeySeat::from_wire (src/pir/session.rs:592-624): y_body is rows x ring_dim
 * coefficient-form key rows; num_to_pack is the key's gamma. */
int inspire_seat_create(inspire_ctx *ctx, const uint64_t *y_body, uint32_t rows, uint32_t num_to_pack,
                        inspire_seat **out, char *err, size_t err_cap);
size_t inspire_seat_resident_bytes(const inspire_seat *seat);
void inspire_seat_destroy(inspire_seat *seat);

/* respond_seeded_batch (src/pir/respond.rs:1189-1202) under one seat: query k is
 * shard_ids[k], the 32-byte seed of its RGSW row (seeds + 32k, SeededRlweCiphertext,
 * src/rlwe/types.rs:149-166) and that row's b (b + k*ring_dim). Outputs per query:
 * a_out + k*ring_dim receives a_hat, b_prefix_out + k*gamma the first gamma
 * coefficients of b (the packed response's prefix, src/pir/respond.rs:132-160).
 * switch_modulus = 0 returns residues mod q; otherwise each coefficient is switched
 * as mod_switch_poly (src/pir/mod_switch.rs:148-175) does, for any target below 2^46
 * (the 36- and 45-bit rungs, MOD_SWITCH_TARGET_36BIT/45BIT, mod_switch.rs:33,47). The
 * switch is mod_switch_response_unchecked (:188); the noise gate of the checked form
 * (check_mod_switch_noise_budget, :111-144) is a host f64 check made once per parameter
 * set by the caller. A refused query fails the
 * whole batch with no partial output (src/pir/respond.rs:1185-1188). Host buffers. */
int inspire_respond_batch(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids,
                          const uint8_t *seeds, const uint64_t *b, size_t count, uint64_t switch_modulus,
                          uint64_t *a_out, uint64_t *b_prefix_out, char *err, size_t err_cap);

/* inspire_respond_batch for unseeded queries (ClientQuery, src/pir/query.rs:103-117):
 * a + k*ring_dim is query k's expanded RGSW row a instead of a seed, as
 * respond_inspiring_cached answers it (src/pir/respond.rs:1022-1041). */
int inspire_respond_batch_expanded(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids,
                                   const uint64_t *a, const uint64_t *b, size_t count, uint64_t switch_modulus,
                                   uint64_t *a_out, uint64_t *b_prefix_out, char *err, size_t err_cap);

/* Device-resident batches for serving loops. A workspace holds the scratch of one
 * in-flight batch of up to max_count queries. */
int inspire_workspace_create(inspire_ctx *ctx, size_t max_count, inspire_workspace **out, char *err,
                             size_t err_cap);
void inspire_workspace_destroy(inspire_workspace *ws);
/* inspire_respond_batch with every buffer in device memory, enqueued on `stream`
 * (a hipStream_t, NULL = default stream) and not waited for. Every shard named must
 * already hold a store; the ids are read on the device, and a query whose id names no shard
 * with a store gets status -1 and no output. Per-query status lands in the workspace: read
 * it with inspire_workspace_status after the stream completes. */
int inspire_respond_batch_device(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                 const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                 size_t count, uint64_t switch_modulus, uint64_t *d_a_out,
                                 uint64_t *d_b_prefix_out, void *stream, char *err, size_t err_cap);
/* Destroying a shard, seat or workspace waits for the device launches of its context enqueued
 * before the call (their fence events), not for the whole device; a launch enqueued afterwards
 * that names a destroyed shard gets status -1. inspire_respond_batch_device_noalloc records no
 * fence event: nothing it reads may be destroyed while it runs. */
/* inspire_respond_batch_device, then waits: kernel_ms[0] receives the DB-scan stage's
 * duration (grouping included) and kernel_ms[1] the recursion kernel's, from HIP events recorded on
 * `stream` around each launch. For measurement. */
int inspire_respond_batch_device_timed(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                       const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                       size_t count, uint64_t switch_modulus, uint64_t *d_a_out,
                                       uint64_t *d_b_prefix_out, void *stream, float kernel_ms[2], char *err,
                                       size_t err_cap);
/* inspire_respond_batch_device without a workspace, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference. It calls no hipMalloc, hipFree or allocating library
 * routine: the DB scan's partial sums live in a library-owned device array of
 * INSPIRE_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), and the per-query
 * status (as inspire_workspace_status reports it) goes to the caller's d_status (count int32).
 * A batch needing more scratch (count * 16 * gamma * 32 bytes at ring 2048: 256 queries at
 * gamma 256) is refused. Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other threads, other streams, other contexts) race and must be serialised by the
 * caller. A shard table that grew since the last launch is refused until one allocating call
 * (any respond or store build) has refreshed it. */
#define INSPIRE_NOALLOC_SCRATCH_BYTES (32u << 20)
int inspire_respond_batch_device_noalloc(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *d_shard_ids,
                                         const uint8_t *d_seeds, const uint64_t *d_b, size_t count,
                                         uint64_t switch_modulus, uint64_t *d_a_out, uint64_t *d_b_prefix_out,
                                         int32_t *d_status, void *stream, char *err, size_t err_cap);
/* Kernel trace (measurement; no counterpart in the reference): arms ws for its next
 * `launches` device-buffer launches (at most 2^20; 0 disarms), each of which then records HIP
 * events on its own stream around its shard-grouping, DB-scan and recursion kernels without
 * waiting, so a timed loop keeps its pipelining. inspire_workspace_kernel_times waits for every
 * traced launch, writes the durations (ms) of up to cap of them to group_ms (0 for a launch
 * that is not grouped), scan_ms and chain_ms, their number to *count, and disarms. Like every
 * use of a workspace, one caller at a time. */
int inspire_workspace_trace(inspire_workspace *ws, size_t launches, char *err, size_t err_cap);
int inspire_workspace_kernel_times(inspire_workspace *ws, float *group_ms, float *scan_ms, float *chain_ms,
                                   size_t cap, size_t *count, char *err, size_t err_cap);
/* Copies count statuses (0 = answered, c+1 = column c failed its store tag, -1 = the id names
 * no shard with a store). */
int inspire_workspace_status(inspire_workspace *ws, size_t count, int32_t *status, char *err, size_t err_cap);

/* ServerResponse::to_binary bytes straight from the device (respond.rs:387, the bincode
 * layout of respond.rs:111-176 with pack_coefficients_tight, poly.rs:166-225): the packing
 * kernel's epilogue switches the modulus (when asked) and packs a_hat and the b prefix at
 * the modulus' bit width into the response body. inspire_response_wire_bytes gives the
 * length (10,446 B at gamma 256 switched to 36 bits; 17,358 B unswitched).
 * _device: responses at d_wire + k * wire_stride; d_wire 4-byte aligned, wire_stride a
 * multiple of 4 and at least the length rounded up to 4 (the pad bytes are written as 0).
 * Host form: responses at wire_out + k * wire_stride, wire_stride >= the length. */
size_t inspire_response_wire_bytes(const inspire_ctx *ctx, uint64_t switch_modulus);
int inspire_respond_batch_device_wire(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                      const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                      size_t count, uint64_t switch_modulus, uint8_t *d_wire, size_t wire_stride,
                                      void *stream, char *err, size_t err_cap);
int inspire_respond_batch_wire(inspire_ctx *ctx, const inspire_seat *seat, const uint32_t *shard_ids,
                               const uint8_t *seeds, const uint64_t *b, size_t count, uint64_t switch_modulus,
                               uint8_t *wire_out, size_t wire_stride, char *err, size_t err_cap);

/* The respond pipeline of inspire_respond_batch_device (DB scan + packing chain) captured
 * once into a hipGraph over fixed device buffers: a serving loop rewrites the input
 * buffers and replays it with one launch. The graph holds kernel nodes only. It stays
 * valid while the context's set of shards is unchanged. No reference counterpart
 * (results are those of inspire_respond_batch_device). */
typedef struct inspire_graph inspire_graph;
int inspire_respond_graph_create(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws,
                                 const uint32_t *d_shard_ids, const uint8_t *d_seeds, const uint64_t *d_b,
                                 size_t count, uint64_t switch_modulus, uint64_t *d_a_out, uint64_t *d_b_prefix_out,
                                 inspire_graph **out, char *err, size_t err_cap);
int inspire_graph_launch(inspire_graph *graph, void *stream, char *err, size_t err_cap);
void inspire_graph_destroy(inspire_graph *graph);

/* Launch-shape autotuner (no reference counterpart; results are unchanged): times the
 * respond pipeline on synthetic queries of the real shape (count queries over the
 * context's stored shards) at 1, 2 and 3 queries per CU, keeps the fastest shape in the
 * context, and reports the three times in kernel_ms and the choice in *chosen. */
int inspire_ctx_autotune(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws, size_t count,
                         float kernel_ms[3], uint32_t *chosen, char *err, size_t err_cap);
/* DB-scan tile autotuner (no reference counterpart; results are unchanged): at ring 2048
 * and up, times the grouping and scan kernels on synthetic queries (count queries at
 * uniform shard ids over the context's stored shards) with the scan split into 8, 16 and
 * 32 row blocks per shard, keeps the fastest in the context (16 until tuned; the next
 * launch of every workspace and serving slot uses it), and reports the three times in
 * scan_ms and the choice in *chosen. Below ring 2048 the scan has one split: *chosen = 1,
 * scan_ms zeros. Fewer splits suit launches that name many shards, more suit small
 * launches (more workgroups per shard). */
int inspire_ctx_autotune_scan(inspire_ctx *ctx, const inspire_seat *seat, inspire_workspace *ws, size_t count,
                              float scan_ms[3], uint32_t *chosen, char *err, size_t err_cap);

/* Continuous-batching server over one context (the serving side of
 * respond_seeded_batch with registered keys, src/pir/respond.rs:1189, and of
 * ServerSessionStore, src/pir/session.rs:700-896). A worker thread merges queued
 * requests from any seats into device launches of up to max_launch_queries queries
 * while the previous launch runs. hbm_budget_bytes caps the seat bytes admitted
 * (0 = limited by free device memory only). */
typedef struct inspire_server inspire_server;
int inspire_server_create(inspire_ctx *ctx, size_t max_launch_queries, size_t hbm_budget_bytes,
                          inspire_server **out, char *err, size_t err_cap);
void inspire_server_destroy(inspire_server *server);
/* ServerSessionStore::register: seat a key in HBM; refused when it does not fit. */
int inspire_server_register_seat(inspire_server *server, const uint64_t *y_body, uint32_t rows,
                                 uint32_t num_to_pack, uint64_t *handle_out, char *err, size_t err_cap);
int inspire_server_unregister_seat(inspire_server *server, uint64_t handle, char *err, size_t err_cap);
/* Queue one batch (count seeded queries) under a registered seat; returns a ticket. */
int inspire_server_submit(inspire_server *server, uint64_t handle, const uint32_t *shard_ids,
                          const uint8_t *seeds, const uint64_t *b, size_t count, uint64_t switch_modulus,
                          uint64_t *ticket_out, char *err, size_t err_cap);
/* Block until the ticket's batch is answered; outputs as inspire_respond_batch.
 * latency_ms receives submit-to-completion time. */
int inspire_server_wait(inspire_server *server, uint64_t ticket, uint64_t *a_out, uint64_t *b_prefix_out,
                        double *latency_ms, char *err, size_t err_cap);
/* Launches and queries answered so far; resident_bytes = device bytes of the seated keys (each
 * key row word with its Shoup companion: 2 x rows x ring_dim x 8 bytes per seat, twice the
 * reference's PackingKeySeat::payload_bytes, src/pir/session.rs:630-636); seats = keys seated. */
int inspire_server_stats(inspire_server *server, uint64_t *launches, uint64_t *queries,
                         uint64_t *resident_bytes, uint64_t *seats);

/* ------------------------------------------------------------------------------
 * Special-modulus set (src/special_modulus): q < 2^29 and a special prime p < 2^31 on
 * 32-bit words, 4-bit records, one step-1 table per shard. Replaces
 * special_modulus::Server::{new, seat, table, respond} (src/special_modulus/server.rs:
 * 145-279). Coefficient arrays are u32 in [0, m). The wire codecs (wire.rs) and the
 * client stay on the host.
 * ---------------------------------------------------------------------------- */
#define INSPIRE_SM_Q 536813569ULL             /* src/special_modulus/mod.rs:49 */
#define INSPIRE_SM_P 2147389441ULL            /* src/special_modulus/mod.rs:52 */
#define INSPIRE_SM_KEY_MASK_VERSION 2         /* src/special_modulus/mod.rs:58 */

typedef struct inspire_sm_server inspire_sm_server;
typedef struct inspire_sm_table inspire_sm_table;
typedef struct inspire_sm_seat inspire_sm_seat;

/* special_modulus::Params (src/special_modulus/mod.rs:62-97). */
typedef struct {
    uint32_t ring_dim;          /* 2048 served; the gfx950 kernels run 16..8192 */
    uint32_t gamma;             /* 512 served */
    uint64_t q;                 /* INSPIRE_SM_Q served; any prime below 2^31 that validate accepts */
    uint64_t special_prime;     /* INSPIRE_SM_P served */
    uint32_t cell_bytes;        /* gamma / 2 */
    uint32_t record_bytes;      /* 246 served */
    double sigma;               /* 6.4 served */
    uint16_t key_mask_version;  /* INSPIRE_SM_KEY_MASK_VERSION */
} inspire_sm_params;

/* Params::validate (mod.rs:100-165), same messages. */
int inspire_sm_params_validate(const inspire_sm_params *params, char *err, size_t err_cap);
/* Server::new / with_kernel (server.rs:145-166): validates, then builds both primes'
 * transform tables on the device. */
int inspire_sm_server_create(const inspire_sm_params *params, inspire_sm_server **out, char *err,
                             size_t err_cap);
void inspire_sm_server_destroy(inspire_sm_server *server);
/* Server::table (server.rs:205-250): len bytes of cells, cell_bytes each, row 0 first.
 * Keeps the cells beside the step-1 table in HBM, so a response is always computed from
 * the cells its table was built from. */
int inspire_sm_table_create(inspire_sm_server *server, const uint8_t *cells, size_t len,
                            inspire_sm_table **out, char *err, size_t err_cap);
uint32_t inspire_sm_table_rows(const inspire_sm_table *table);          /* ShardTable::rows */
size_t inspire_sm_table_resident_bytes(const inspire_sm_table *table);  /* device bytes held */
void inspire_sm_table_destroy(inspire_sm_table *table);
/* Server::seat (server.rs:177-203): the key's two limbs (coefficient form, ring_dim words
 * each) and the seed its mask expands from, as PackingKey::from_bytes decoded them. */
int inspire_sm_seat_create(inspire_sm_server *server, const uint8_t mask_seed[32], const uint32_t *key_q,
                           const uint32_t *key_p, inspire_sm_seat **out, char *err, size_t err_cap);
size_t inspire_sm_seat_resident_bytes(const inspire_sm_seat *seat);
void inspire_sm_seat_destroy(inspire_sm_seat *seat);
/* Server::respond (server.rs:252-279): one query (32-byte seed, ring_dim body words);
 * writes the mask (ring_dim words) and the body (2 * record_bytes words). */
int inspire_sm_respond(inspire_sm_server *server, const inspire_sm_table *table, const inspire_sm_seat *seat,
                       const uint8_t seed[32], const uint32_t *body, uint32_t *mask_out, uint32_t *body_out,
                       char *err, size_t err_cap);
/* count queries in one launch, each with its own table and seat (the loop a caller of
 * Server::respond runs over a batch); arrays are query-major. */
int inspire_sm_respond_batch(inspire_sm_server *server, size_t count, const inspire_sm_table *const *tables,
                             const inspire_sm_seat *const *seats, const uint8_t *seeds, const uint32_t *bodies,
                             uint32_t *mask_out, uint32_t *body_out, char *err, size_t err_cap);
/* The same with device-resident seeds, bodies and outputs, launched on stream (a
 * hipStream_t, NULL for the default stream). Bodies are taken as residues: the host
 * entry points refuse one that is not. With kernel_ms non-NULL the call waits for the
 * launch and returns its duration from events on the stream. */
int inspire_sm_respond_batch_device(inspire_sm_server *server, size_t count,
                                    const inspire_sm_table *const *tables, const inspire_sm_seat *const *seats,
                                    const uint8_t *d_seeds, const uint32_t *d_bodies, uint32_t *d_mask_out,
                                    uint32_t *d_body_out, void *stream, float *kernel_ms, char *err,
                                    size_t err_cap);
/* Length of one Response::to_bytes (wire.rs:125-131): the mask then the body, each packed
 * LSB-first at the bit width of q - 1 (9,208 B at ring2048_cell256). */
size_t inspire_sm_response_wire_bytes(const inspire_sm_server *server);
/* inspire_sm_respond_batch returning each query's Response::to_bytes, packed by the kernel:
 * query k's bytes at wire_out + k * wire_stride (wire_stride >= the response length). */
int inspire_sm_respond_batch_wire(inspire_sm_server *server, size_t count, const inspire_sm_table *const *tables,
                                  const inspire_sm_seat *const *seats, const uint8_t *seeds, const uint32_t *bodies,
                                  uint8_t *wire_out, size_t wire_stride, char *err, size_t err_cap);
/* The same on device buffers, launched on stream. d_scratch holds count * ring_dim words that
 * the chain uses while it runs (the mask output's role in inspire_sm_respond_batch_device);
 * wire_stride is a multiple of 4 at least the response length rounded up to 4. */
int inspire_sm_respond_batch_device_wire(inspire_sm_server *server, size_t count,
                                         const inspire_sm_table *const *tables, const inspire_sm_seat *const *seats,
                                         const uint8_t *d_seeds, const uint32_t *d_bodies, uint32_t *d_scratch,
                                         uint8_t *d_wire_out, size_t wire_stride, void *stream, char *err,
                                         size_t err_cap);
/* Kernel trace (measurement; no counterpart in the reference): the server's next `launches`
 * device-buffer launches (the two entry points above; at most 2^20, 0 disarms) record HIP
 * events around the respond kernel on their own streams without waiting;
 * inspire_sm_server_kernel_times waits for the last one, writes up to cap durations (ms) and
 * their number, and disarms. */
int inspire_sm_server_trace(inspire_sm_server *server, size_t launches, char *err, size_t err_cap);
int inspire_sm_server_kernel_times(inspire_sm_server *server, float *kernel_ms, size_t cap, size_t *count,
                                   char *err, size_t err_cap);

/* Continuous-batching queue over one special-modulus server: requests (a batch of
 * queries under one seated key, each query naming its own table) from any thread are
 * merged into launches of up to max_launch_queries while the previous launch runs; every
 * query's result equals its Server::respond alone. Keys are seated by handle and admitted
 * against hbm_budget_bytes (0 = free device memory only). Tables must outlive the
 * requests that name them. No counterpart type in the reference, which answers each
 * query on the caller's thread (server.rs:252-279). */
typedef struct inspire_sm_batcher inspire_sm_batcher;
int inspire_sm_batcher_create(inspire_sm_server *server, size_t max_launch_queries, size_t hbm_budget_bytes,
                              inspire_sm_batcher **out, char *err, size_t err_cap);
void inspire_sm_batcher_destroy(inspire_sm_batcher *batcher);
/* Server::seat for a key kept by handle. */
int inspire_sm_batcher_register_key(inspire_sm_batcher *batcher, const uint8_t mask_seed[32], const uint32_t *key_q,
                                    const uint32_t *key_p, uint64_t *handle_out, char *err, size_t err_cap);
int inspire_sm_batcher_unregister_key(inspire_sm_batcher *batcher, uint64_t handle, char *err, size_t err_cap);
int inspire_sm_batcher_submit(inspire_sm_batcher *batcher, uint64_t handle, const inspire_sm_table *const *tables,
                              const uint8_t *seeds, const uint32_t *bodies, size_t count, uint64_t *ticket_out,
                              char *err, size_t err_cap);
/* Blocks until the ticket is answered; outputs as inspire_sm_respond_batch, latency_ms =
 * submit to completion. */
int inspire_sm_batcher_wait(inspire_sm_batcher *batcher, uint64_t ticket, uint32_t *mask_out, uint32_t *body_out,
                            double *latency_ms, char *err, size_t err_cap);
int inspire_sm_batcher_stats(inspire_sm_batcher *batcher, uint64_t *launches, uint64_t *queries,
                             uint64_t *resident_bytes);

#ifdef __cplusplus
}
#endif
#endif
