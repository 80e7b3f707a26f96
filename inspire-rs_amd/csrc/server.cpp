// Continuous-batching respond server on one GPU.
//
// The reference answers a SeededQueryBatch sequentially on the calling thread
// (respond_seeded_batch, src/pir/respond.rs:1189-1202) and resolves a registered key
// through ServerSessionStore (src/pir/session.rs:700-896). Here every request (one
// batch under one registered seat) is queued; a worker thread merges whatever is
// queued, across seats and batch sizes, into one device launch of up to
// max_launch queries (each query carries its own seat pointer), and keeps up to two
// launches on the GPU, each on its own stream so that they overlap (the next launch's scan
// and first chains run while the previous one's last chains drain): requests that arrive
// while one launch runs form the next. A
// request's queries are never split across launches, and every query's result is
// the same bytes it gets alone (the kernels read nothing across queries).
//
// Seats live in HBM by handle and are admitted against free device memory and an
// optional byte budget; the resident byte count follows PackingKeySeat::resident_bytes
// (src/pir/session.rs:632-652) for the device copy.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "runtime.hpp"

namespace {

using clock_type = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    std::shared_ptr<inspire_seat> seat;
    uint64_t switch_modulus = 0;
    std::vector<uint32_t> ids;
    std::vector<uint8_t> seeds;
    std::vector<uint64_t> b;
    clock_type::time_point submitted;
};

struct Result {
    bool done = false;
    int rc = INSPIRE_OK;
    std::string err;
    std::vector<uint64_t> a, bp;
    double latency_ms = 0;
};

// Pinned staging and device buffers of one launch slot.
struct Slot {
    size_t cap = 0;
    uint32_t *h_ids = nullptr;
    uint8_t *h_seeds = nullptr;
    uint64_t *h_b = nullptr;
    const Pair **h_seats = nullptr;
    uint64_t *h_a = nullptr, *h_bp = nullptr;
    int32_t *h_status = nullptr;
    DevBuf d_ids, d_seeds, d_b, d_seats, d_a, d_bp, scan_part, status, group;
    // inputs copied (copy-in stream), kernels done (compute stream), outputs copied (copy-out
    // stream): the last one completes the slot
    hipEvent_t in_ev = nullptr, kern_ev = nullptr, done_ev = nullptr;
    // the slot's kernels run here: the two slots' launches overlap on the device, so one
    // launch's scan and its first chains fill the CUs the other's last chains leave idle
    hipStream_t compute = nullptr;
    std::vector<Request> reqs;
    size_t count = 0;
    bool busy = false;
    ~Slot() {
        (void)hipHostFree(h_ids); (void)hipHostFree(h_seeds); (void)hipHostFree(h_b);
        (void)hipHostFree(h_seats); (void)hipHostFree(h_a); (void)hipHostFree(h_bp); (void)hipHostFree(h_status);
        if (in_ev) (void)hipEventDestroy(in_ev);
        if (kern_ev) (void)hipEventDestroy(kern_ev);
        if (done_ev) (void)hipEventDestroy(done_ev);
    }
};

}  // namespace

struct inspire_server {
    inspire_ctx *ctx = nullptr;
    size_t max_launch = 0;
    size_t hbm_budget = 0;     // 0 = free device memory only
    size_t resident = 0;       // seat bytes admitted
    // Kernels run on each slot's compute stream; a slot's inputs go up on `copy_in` and its
    // outputs come back on `copy_out`, so the copies of one launch overlap the kernels of its
    // neighbours. All are non-blocking streams at the default priority.
    hipStream_t copy_in = nullptr, copy_out = nullptr;
    Slot slots[2];

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::map<uint64_t, Result> results;
    std::map<uint64_t, std::shared_ptr<inspire_seat>> seats;
    uint64_t next_ticket = 1, next_handle = 1;
    bool stopping = false;
    std::thread worker;
    std::atomic<uint64_t> launches{0}, queries{0};

    void run();
    int launch(Slot &s, char *err, size_t err_cap);
    void complete(Slot &s);
};

namespace {

int alloc_slot(inspire_server *sv, Slot &s, char *err, size_t err_cap) {
    const size_t c = sv->max_launch, n = sv->ctx->n, g = sv->ctx->gamma;
    HIP_TRY(hipHostMalloc((void **)&s.h_ids, c * 4, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.h_seeds, c * 32, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.h_b, c * n * 8, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.h_seats, c * sizeof(void *), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.h_a, c * n * 8, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.h_bp, c * g * 8, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.h_status, c * 4, hipHostMallocDefault));
    HIP_TRY(s.d_ids.alloc(c * 4));
    HIP_TRY(s.d_seeds.alloc(c * 32));
    HIP_TRY(s.d_b.alloc(c * n * 8));
    HIP_TRY(s.d_seats.alloc(c * sizeof(void *)));
    HIP_TRY(s.d_a.alloc(c * n * 8));
    HIP_TRY(s.d_bp.alloc(c * g * 8));
    HIP_TRY(s.scan_part.alloc(c * sv->ctx->ks_cap() * g * 4 * 8));
    HIP_TRY(s.status.alloc(c * 4));
    const size_t gc = c < (size_t)kGroupMax ? c : (size_t)kGroupMax;
    HIP_TRY(s.group.alloc((2 * gc + 2) * 4));
    HIP_TRY(hipEventCreateWithFlags(&s.in_ev, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&s.kern_ev, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming));
    s.cap = c;
    return INSPIRE_OK;
}

}  // namespace

int inspire_server::launch(Slot &s, char *err, size_t err_cap) {
    const size_t n = ctx->n, g = ctx->gamma;
    size_t q = 0;
    for (const Request &r : s.reqs) {
        const size_t c = r.ids.size();
        memcpy(s.h_ids + q, r.ids.data(), c * 4);
        memcpy(s.h_seeds + q * 32, r.seeds.data(), c * 32);
        memcpy(s.h_b + q * n, r.b.data(), c * n * 8);
        for (size_t k = 0; k < c; k++) s.h_seats[q + k] = r.seat->yprime.as<Pair>();
        q += c;
    }
    s.count = q;
    // the shard table and launch shape as of this launch, read under the context's lock, which is
    // held until the kernels are enqueued with a fence event behind them (a shard or seat
    // destroyed meanwhile waits for this launch)
    std::lock_guard<std::mutex> lk(ctx->mu);
    int rc = sync_table(ctx, err, err_cap);
    if (rc) return rc;
    const CtxDev dev = ctx->dev();
    // The slot's previous launch has completed (its done_ev, after its output copies), so its
    // buffers are free; the copies may run while the other slot's kernels do.
    HIP_TRY(hipMemcpyAsync(s.d_ids.p, s.h_ids, q * 4, hipMemcpyHostToDevice, copy_in));
    HIP_TRY(hipMemcpyAsync(s.d_seeds.p, s.h_seeds, q * 32, hipMemcpyHostToDevice, copy_in));
    HIP_TRY(hipMemcpyAsync(s.d_b.p, s.h_b, q * n * 8, hipMemcpyHostToDevice, copy_in));
    HIP_TRY(hipMemcpyAsync(s.d_seats.p, s.h_seats, q * sizeof(void *), hipMemcpyHostToDevice, copy_in));
    HIP_TRY(hipEventRecord(s.in_ev, copy_in));
    HIP_TRY(hipStreamWaitEvent(s.compute, s.in_ev, 0));
    // seats registered since their kernels were last seen complete: this launch waits for them
    for (const Request &r : s.reqs) {
        inspire_seat *st = r.seat.get();
        if (st->settled) continue;
        if (hipEventQuery(st->ready) == hipSuccess)
            st->settled = true;
        else
            HIP_TRY(hipStreamWaitEvent(s.compute, st->ready, 0));
    }
    BatchDev bd;
    bd.count = (int)q;
    bd.shard_ids = s.d_ids.as<uint32_t>();
    bd.seeds = s.d_seeds.as<uint8_t>();
    bd.a_in = nullptr;
    bd.b = s.d_b.as<uint64_t>();
    bd.yprime = nullptr;
    bd.yprimes = s.d_seats.as<const Pair *>();
    bd.scan_part = s.scan_part.as<uint64_t>();
    bd.ks = ctx->ks();
    bd.a_out = s.d_a.as<uint64_t>();
    bd.b_out = s.d_bp.as<uint64_t>();
    bd.status = s.status.as<int32_t>();
    bd.switch_modulus = s.reqs.front().switch_modulus;
    const size_t gc = s.cap < (size_t)kGroupMax ? s.cap : (size_t)kGroupMax;
    if (q <= gc) {  // shard grouping of the launch (see BatchDev::order)
        uint32_t *g = s.group.as<uint32_t>();
        bd.order = g;
        bd.first = g + gc;
        bd.ngroups = g + 2 * gc + 1;
    }
    HIP_TRY(launch_respond(dev, bd, s.compute));
    HIP_TRY(ctx->fence.note(s.compute));
    HIP_TRY(hipEventRecord(s.kern_ev, s.compute));
    HIP_TRY(hipStreamWaitEvent(copy_out, s.kern_ev, 0));
    HIP_TRY(hipMemcpyAsync(s.h_a, s.d_a.p, q * n * 8, hipMemcpyDeviceToHost, copy_out));
    HIP_TRY(hipMemcpyAsync(s.h_bp, s.d_bp.p, q * g * 8, hipMemcpyDeviceToHost, copy_out));
    HIP_TRY(hipMemcpyAsync(s.h_status, s.status.p, q * 4, hipMemcpyDeviceToHost, copy_out));
    HIP_TRY(hipEventRecord(s.done_ev, copy_out));
    launches++;
    queries += q;
    return INSPIRE_OK;
}

void inspire_server::complete(Slot &s) {
    hipError_t e = hipEventSynchronize(s.done_ev);
    const size_t n = ctx->n, g = ctx->gamma;
    const auto now = clock_type::now();
    std::lock_guard<std::mutex> lk(mu);
    size_t q = 0;
    for (Request &r : s.reqs) {
        Result &res = results[r.ticket];
        const size_t c = r.ids.size();
        res.latency_ms = std::chrono::duration<double, std::milli>(now - r.submitted).count();
        if (e != hipSuccess) {
            res.rc = INSPIRE_ERR_DEVICE;
            res.err = std::string("HIP failure completing a launch: ") + hipGetErrorString(e);
        } else {
            for (size_t k = 0; k < c && res.rc == INSPIRE_OK; k++)
                if (s.h_status[q + k] < 0) {
                    res.rc = INSPIRE_ERR_STORE;
                    res.err = "shard " + std::to_string(r.ids[k]) + " has no packing store on the device";
                } else if (s.h_status[q + k] != 0) {
                    char buf[320];
                    snprintf(buf, sizeof buf,
                             "shard %u: packing store refused: column %d is not the column the store was built "
                             "from. The shard changed after the build; rebuild the store before serving it",
                             r.ids[k], s.h_status[q + k] - 1);
                    res.rc = INSPIRE_ERR_STORE;
                    res.err = buf;
                }
            if (res.rc == INSPIRE_OK) {
                res.a.assign(s.h_a + q * n, s.h_a + (q + c) * n);
                res.bp.assign(s.h_bp + q * g, s.h_bp + (q + c) * g);
            }
        }
        res.done = true;
        q += c;
    }
    s.reqs.clear();
    s.busy = false;
    cv_done.notify_all();
}

void inspire_server::run() {
    int next = 0;
    for (;;) {
        Slot &s = slots[next];
        if (s.busy) complete(s);  // the slot's previous launch has to finish first
        {
            std::unique_lock<std::mutex> lk(mu);
            Slot &other = slots[next ^ 1];
            while (pending.empty() && !stopping) {
                if (other.busy) {
                    lk.unlock();
                    complete(other);
                    lk.lock();
                    continue;
                }
                cv_work.wait(lk);
            }
            if (pending.empty() && stopping) break;
            // merge queued requests (FIFO) under one switch modulus up to max_launch queries
            size_t q = 0;
            const uint64_t mod = pending.front().switch_modulus;
            while (!pending.empty() && pending.front().switch_modulus == mod &&
                   q + pending.front().ids.size() <= max_launch) {
                q += pending.front().ids.size();
                s.reqs.push_back(std::move(pending.front()));
                pending.pop_front();
            }
        }
        char err[512] = {0};
        int rc = launch(s, err, sizeof err);
        if (rc != INSPIRE_OK) {
            // a partly enqueued launch may still read or write the slot's buffers
            (void)hipStreamSynchronize(copy_in);
            (void)hipStreamSynchronize(s.compute);
            (void)hipStreamSynchronize(copy_out);
            std::lock_guard<std::mutex> lk(mu);
            for (Request &r : s.reqs) {
                Result &res = results[r.ticket];
                res.rc = rc;
                res.err = err;
                res.done = true;
            }
            s.reqs.clear();
            cv_done.notify_all();
            continue;
        }
        s.busy = true;
        next ^= 1;
    }
    for (Slot &s : slots)
        if (s.busy) complete(s);
}

extern "C" {

int inspire_server_create(inspire_ctx *ctx, size_t max_launch_queries, size_t hbm_budget_bytes, inspire_server **out,
                          char *err, size_t err_cap) {
    if (!ctx || !out || max_launch_queries == 0 || max_launch_queries > INSPIRE_MAX_BATCH) {
        set_err(err, err_cap, "server needs a context and a launch size of 1 to 4194304 queries");
        return INSPIRE_ERR_INVALID;
    }
    if (!ctx->has_crs) {
        set_err(err, err_cap, kNoCrs);
        return INSPIRE_ERR_INVALID;
    }
    inspire_server *sv = new (std::nothrow) inspire_server();
    if (!sv) { set_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    sv->ctx = ctx;
    sv->max_launch = max_launch_queries;
    sv->hbm_budget = hbm_budget_bytes;
    auto drop_streams = [sv] {
        for (hipStream_t st : {sv->slots[0].compute, sv->slots[1].compute, sv->copy_in, sv->copy_out})
            if (st) (void)hipStreamDestroy(st);
    };
    for (hipStream_t *st : {&sv->slots[0].compute, &sv->slots[1].compute, &sv->copy_in, &sv->copy_out}) {
        hipError_t e = hipStreamCreateWithFlags(st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_err(err, err_cap, "HIP failure creating the server streams: %s", hipGetErrorString(e));
            drop_streams();
            delete sv;
            return INSPIRE_ERR_DEVICE;
        }
    }
    for (Slot &s : sv->slots) {
        int rc = alloc_slot(sv, s, err, err_cap);
        if (rc) {
            drop_streams();
            delete sv;
            return rc;
        }
    }
    sv->worker = std::thread([sv] { sv->run(); });
    *out = sv;
    return INSPIRE_OK;
}

void inspire_server_destroy(inspire_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
    }
    sv->cv_work.notify_all();
    sv->worker.join();
    for (hipStream_t st : {sv->copy_in, sv->slots[0].compute, sv->slots[1].compute, sv->copy_out}) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete sv;
}

int inspire_server_register_seat(inspire_server *sv, const uint64_t *y_body, uint32_t rows, uint32_t num_to_pack,
                                 uint64_t *handle_out, char *err, size_t err_cap) {
    if (!sv || !handle_out) { set_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    const size_t bytes = (size_t)rows * sv->ctx->n * sizeof(Pair);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        // keep headroom of two launches' worth of buffers plus 256 MiB
        const size_t headroom = (size_t)256 << 20;
        if ((sv->hbm_budget && sv->resident + bytes > sv->hbm_budget) || free_b < bytes + headroom) {
            set_err(err, err_cap,
                    "ServerSessionStore admission refused: seating %zu more bytes would exceed the device budget "
                    "(%zu of %zu bytes resident, %zu bytes free on the device)",
                    bytes, sv->resident, sv->hbm_budget, free_b);
            return INSPIRE_ERR_INVALID;
        }
    }
    // the seat's kernels are enqueued on the context's stream and not waited for: the launch
    // that first carries one of its requests waits for them on the device (launch())
    inspire_seat *seat = nullptr;
    int rc = seat_create_enqueue(sv->ctx, y_body, rows, num_to_pack, &seat, err, err_cap);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(sv->mu);
    const uint64_t h = sv->next_handle++;
    sv->seats[h] = std::shared_ptr<inspire_seat>(seat, [](inspire_seat *p) { inspire_seat_destroy(p); });
    sv->resident += bytes;
    *handle_out = h;
    return INSPIRE_OK;
}

int inspire_server_unregister_seat(inspire_server *sv, uint64_t handle, char *err, size_t err_cap) {
    if (!sv) { set_err(err, err_cap, "null server"); return INSPIRE_ERR_INVALID; }
    // queued requests keep the seat alive until they complete; otherwise it is destroyed when
    // `dropped` goes, after the lock (the destroy waits for the launches enqueued so far, and
    // the worker must not wait behind it)
    std::shared_ptr<inspire_seat> dropped;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->seats.find(handle);
    if (it == sv->seats.end()) {
        set_err(err, err_cap, "ServerSessionStore has no entry for session_handle ServerSessionHandle(%llu)",
                (unsigned long long)handle);
        return INSPIRE_ERR_INVALID;
    }
    sv->resident -= (size_t)it->second->rows * sv->ctx->n * sizeof(Pair);
    dropped = std::move(it->second);
    sv->seats.erase(it);
    return INSPIRE_OK;
}

int inspire_server_submit(inspire_server *sv, uint64_t handle, const uint32_t *shard_ids, const uint8_t *seeds,
                          const uint64_t *b, size_t count, uint64_t switch_modulus, uint64_t *ticket_out, char *err,
                          size_t err_cap) {
    if (!sv || !ticket_out || (count && (!shard_ids || !seeds || !b))) {
        set_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    inspire_ctx *ctx = sv->ctx;
    if (count == 0 || count > sv->max_launch) {
        set_err(err, err_cap, "a request carries 1..=%zu queries, got %zu", sv->max_launch, count);
        return INSPIRE_ERR_INVALID;
    }
    if (switch_modulus >= kSwitchModulusLimit) {
        set_err(err, err_cap, "switch modulus %llu is not below 2^46 (the 36- and 45-bit targets are implemented)",
                (unsigned long long)switch_modulus);
        return INSPIRE_ERR_INVALID;
    }
    const size_t n = ctx->n;
    for (size_t k = 0; k < count * n; k++)
        if (b[k] >= kQ) {
            set_err(err, err_cap, "Poly decode refused: limb 0 coefficient %zu is %llu, not canonical for modulus %llu",
                    k % n, (unsigned long long)b[k], (unsigned long long)kQ);
            return INSPIRE_ERR_INVALID;
        }
    {
        // shard checks; a shard without a store gets one built now (the bytes equal the
        // per-query path's, shard_store.rs:1-16)
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (size_t k = 0; k < count; k++) {
            const uint32_t id = shard_ids[k];
            if (id >= ctx->owners.size() || !ctx->owners[id]) {
                set_err(err, err_cap,
                        "respond_seeded_with_seat: shard %u not found at dense position %u; encoded database has %zu "
                        "shards",
                        id, id, ctx->owners.size());
                return INSPIRE_ERR_INVALID;
            }
            if (!ctx->owners[id]->has_store) {
                int rc = build_store(ctx, ctx->owners[id], true, err, err_cap);
                if (rc) return rc;
            }
        }
    }
    Request r;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->seats.find(handle);
        if (it == sv->seats.end()) {
            set_err(err, err_cap, "ServerSessionStore has no entry for session_handle ServerSessionHandle(%llu)",
                    (unsigned long long)handle);
            return INSPIRE_ERR_INVALID;
        }
        r.seat = it->second;
        r.ticket = sv->next_ticket++;
    }
    r.switch_modulus = switch_modulus;
    r.ids.assign(shard_ids, shard_ids + count);
    r.seeds.assign(seeds, seeds + count * 32);
    r.b.assign(b, b + count * n);
    r.submitted = clock_type::now();
    const uint64_t ticket = r.ticket;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->results[ticket] = Result();
        sv->pending.push_back(std::move(r));
    }
    sv->cv_work.notify_one();
    *ticket_out = ticket;
    return INSPIRE_OK;
}

int inspire_server_wait(inspire_server *sv, uint64_t ticket, uint64_t *a_out, uint64_t *b_prefix_out,
                        double *latency_ms, char *err, size_t err_cap) {
    if (!sv) { set_err(err, err_cap, "null server"); return INSPIRE_ERR_INVALID; }
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->results.find(ticket) == sv->results.end()) {
        set_err(err, err_cap, "unknown ticket %llu", (unsigned long long)ticket);
        return INSPIRE_ERR_INVALID;
    }
    // found again on every wake-up: another waiter on the same ticket may have collected it
    auto it = sv->results.end();
    sv->cv_done.wait(lk, [&] {
        it = sv->results.find(ticket);
        return it == sv->results.end() || it->second.done;
    });
    if (it == sv->results.end()) {
        set_err(err, err_cap, "ticket %llu was collected by another wait", (unsigned long long)ticket);
        return INSPIRE_ERR_INVALID;
    }
    Result res = std::move(it->second);
    sv->results.erase(it);
    lk.unlock();
    if (latency_ms) *latency_ms = res.latency_ms;
    if (res.rc != INSPIRE_OK) {
        set_err(err, err_cap, "%s", res.err.c_str());
        return res.rc;
    }
    if (a_out) memcpy(a_out, res.a.data(), res.a.size() * 8);
    if (b_prefix_out) memcpy(b_prefix_out, res.bp.data(), res.bp.size() * 8);
    return INSPIRE_OK;
}

int inspire_server_stats(inspire_server *sv, uint64_t *launches, uint64_t *queries, uint64_t *resident_bytes,
                         uint64_t *seats) {
    if (!sv) return INSPIRE_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (launches) *launches = sv->launches.load();
    if (queries) *queries = sv->queries.load();
    if (resident_bytes) *resident_bytes = sv->resident;
    if (seats) *seats = sv->seats.size();
    return INSPIRE_OK;
}

}  // extern "C"
