// Continuous-batching queue of the special-modulus set on one GPU.
//
// The reference answers each query with Server::respond on the caller's thread
// (src/special_modulus/server.rs:252-279). Here a request is a batch of queries under one
// registered key, each query naming its own step-1 table; requests are queued, and a
// worker thread merges whatever is queued, across keys, tables and batch sizes, into one
// launch of up to max_launch queries while the previous launch runs (two launch slots
// with pinned staging and asynchronous copies). A request's queries are never split
// across launches, and every query's result is the bytes it gets alone: the kernel reads
// nothing across queries.
//
// Keys are seated in HBM by handle and admitted against free device memory and an
// optional byte budget; a seat holds KeySeat::bytes() = 32 KiB at the served size
// (server.rs:85-92, mod.rs:201-203).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <thread>

#include "sm_runtime.hpp"

using namespace inspire;
using namespace inspire_sm;

namespace {

using clock_type = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    std::shared_ptr<inspire_sm_seat> seat;
    std::vector<const inspire_sm_table *> tables;
    std::vector<uint8_t> seeds;
    std::vector<uint32_t> bodies;
    clock_type::time_point submitted;
};

struct Result {
    bool done = false;
    int rc = INSPIRE_OK;
    std::string err;
    std::vector<uint32_t> mask, body;
    double latency_ms = 0;
};

// Pinned staging and device buffers of one launch slot.
struct Slot {
    SmJob *h_jobs = nullptr;
    uint8_t *h_seeds = nullptr;
    uint32_t *h_bodies = nullptr, *h_mask = nullptr, *h_body = nullptr;
    Buf d_jobs, d_seeds, d_bodies, d_mask, d_body;
    // inputs copied (copy-in stream), kernel done (compute stream), outputs copied (copy-out
    // stream): the last one completes the slot
    hipEvent_t in_ev = nullptr, kern_ev = nullptr, done_ev = nullptr;
    // the slot's kernel runs here: the two slots' launches overlap on the device
    hipStream_t compute = nullptr;
    std::vector<Request> reqs;
    bool busy = false;
    ~Slot() {
        (void)hipHostFree(h_jobs); (void)hipHostFree(h_seeds); (void)hipHostFree(h_bodies);
        (void)hipHostFree(h_mask); (void)hipHostFree(h_body);
        if (in_ev) (void)hipEventDestroy(in_ev);
        if (kern_ev) (void)hipEventDestroy(kern_ev);
        if (done_ev) (void)hipEventDestroy(done_ev);
    }
};

}  // namespace

struct inspire_sm_batcher {
    inspire_sm_server *srv = nullptr;
    size_t max_launch = 0, hbm_budget = 0, resident = 0;
    size_t n = 0, nb = 0;
    // The kernel runs on each slot's compute stream; a slot's inputs go up on `copy_in` and
    // its outputs come back on `copy_out`, so one launch's copies overlap its neighbours'
    // kernels. All are non-blocking streams at the default priority.
    hipStream_t copy_in = nullptr, copy_out = nullptr;
    Slot slots[2];

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::map<uint64_t, Result> results;
    std::map<uint64_t, std::shared_ptr<inspire_sm_seat>> seats;
    uint64_t next_ticket = 1, next_handle = 1;
    bool stopping = false;
    std::thread worker;
    std::atomic<uint64_t> launches{0}, queries{0};

    void run();
    int launch(Slot &s, char *err, size_t err_cap);
    void complete(Slot &s);
};

namespace {

int alloc_slot(inspire_sm_batcher *b, Slot &s, char *err, size_t err_cap) {
    const size_t c = b->max_launch, n = b->n, nb = b->nb;
    SM_TRY(hipHostMalloc((void **)&s.h_jobs, c * sizeof(SmJob), hipHostMallocDefault));
    SM_TRY(hipHostMalloc((void **)&s.h_seeds, c * 32, hipHostMallocDefault));
    SM_TRY(hipHostMalloc((void **)&s.h_bodies, c * n * 4, hipHostMallocDefault));
    SM_TRY(hipHostMalloc((void **)&s.h_mask, c * n * 4, hipHostMallocDefault));
    SM_TRY(hipHostMalloc((void **)&s.h_body, c * nb * 4, hipHostMallocDefault));
    SM_TRY(s.d_jobs.alloc(c * sizeof(SmJob)));
    SM_TRY(s.d_seeds.alloc(c * 32));
    SM_TRY(s.d_bodies.alloc(c * n * 4));
    SM_TRY(s.d_mask.alloc(c * n * 4));
    SM_TRY(s.d_body.alloc(c * nb * 4));
    SM_TRY(hipEventCreateWithFlags(&s.in_ev, hipEventDisableTiming));
    SM_TRY(hipEventCreateWithFlags(&s.kern_ev, hipEventDisableTiming));
    SM_TRY(hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming));
    return INSPIRE_OK;
}

}  // namespace

int inspire_sm_batcher::launch(Slot &s, char *err, size_t err_cap) {
    size_t q = 0;
    for (const Request &r : s.reqs) {
        const size_t c = r.tables.size();
        for (size_t k = 0; k < c; k++)
            s.h_jobs[q + k] = SmJob{r.tables[k]->step1.as<uint32_t>(), r.tables[k]->cells.as<uint8_t>(),
                                    r.seat->words.as<uint32_t>(), (int)(q + k)};
        memcpy(s.h_seeds + q * 32, r.seeds.data(), c * 32);
        memcpy(s.h_bodies + q * n, r.bodies.data(), c * n * 4);
        q += c;
    }
    sm_order_jobs(s.h_jobs, q);
    // the slot's previous launch has completed (its done_ev), so its buffers are free
    SM_TRY(hipMemcpyAsync(s.d_jobs.p, s.h_jobs, q * sizeof(SmJob), hipMemcpyHostToDevice, copy_in));
    SM_TRY(hipMemcpyAsync(s.d_seeds.p, s.h_seeds, q * 32, hipMemcpyHostToDevice, copy_in));
    SM_TRY(hipMemcpyAsync(s.d_bodies.p, s.h_bodies, q * n * 4, hipMemcpyHostToDevice, copy_in));
    SM_TRY(hipEventRecord(s.in_ev, copy_in));
    SM_TRY(hipStreamWaitEvent(s.compute, s.in_ev, 0));
    // seats registered since their kernel was last seen complete: this launch waits for them
    for (const Request &r : s.reqs) {
        inspire_sm_seat *st = r.seat.get();
        if (st->settled) continue;
        if (hipEventQuery(st->ready) == hipSuccess)
            st->settled = true;
        else
            SM_TRY(hipStreamWaitEvent(s.compute, st->ready, 0));
    }
    SmBatchDev bd{(int)q, s.d_jobs.as<SmJob>(), s.d_seeds.as<uint8_t>(), s.d_bodies.as<uint32_t>(),
                  s.d_mask.as<uint32_t>(), s.d_body.as<uint32_t>()};
    {
        // a fence event behind the kernel, so destroying a seat or table of the server waits for it
        std::lock_guard<std::mutex> lk(srv->mu);
        SM_TRY(sm_launch_respond(srv->dev, bd, s.compute));
        SM_TRY(srv->fence.note(s.compute));
    }
    SM_TRY(hipEventRecord(s.kern_ev, s.compute));
    SM_TRY(hipStreamWaitEvent(copy_out, s.kern_ev, 0));
    SM_TRY(hipMemcpyAsync(s.h_mask, s.d_mask.p, q * n * 4, hipMemcpyDeviceToHost, copy_out));
    SM_TRY(hipMemcpyAsync(s.h_body, s.d_body.p, q * nb * 4, hipMemcpyDeviceToHost, copy_out));
    SM_TRY(hipEventRecord(s.done_ev, copy_out));
    launches++;
    queries += q;
    return INSPIRE_OK;
}

void inspire_sm_batcher::complete(Slot &s) {
    hipError_t e = hipEventSynchronize(s.done_ev);
    const auto now = clock_type::now();
    std::lock_guard<std::mutex> lk(mu);
    size_t q = 0;
    for (Request &r : s.reqs) {
        Result &res = results[r.ticket];
        const size_t c = r.tables.size();
        res.latency_ms = std::chrono::duration<double, std::milli>(now - r.submitted).count();
        if (e != hipSuccess) {
            res.rc = INSPIRE_ERR_DEVICE;
            res.err = std::string("HIP failure completing a launch: ") + hipGetErrorString(e);
        } else {
            res.mask.assign(s.h_mask + q * n, s.h_mask + (q + c) * n);
            res.body.assign(s.h_body + q * nb, s.h_body + (q + c) * nb);
        }
        res.done = true;
        q += c;
    }
    s.reqs.clear();
    s.busy = false;
    cv_done.notify_all();
}

void inspire_sm_batcher::run() {
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
            // merge queued requests (FIFO) up to max_launch queries
            size_t q = 0;
            while (!pending.empty() && q + pending.front().tables.size() <= max_launch) {
                q += pending.front().tables.size();
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

int inspire_sm_batcher_create(inspire_sm_server *srv, size_t max_launch_queries, size_t hbm_budget_bytes,
                              inspire_sm_batcher **out, char *err, size_t err_cap) {
    if (!srv || !out || max_launch_queries == 0 || max_launch_queries > INSPIRE_MAX_BATCH) {
        sm_err(err, err_cap, "a batcher needs a server and a launch size of 1 to 4194304 queries");
        return INSPIRE_ERR_INVALID;
    }
    inspire_sm_batcher *b = new (std::nothrow) inspire_sm_batcher();
    if (!b) { sm_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    b->srv = srv;
    b->max_launch = max_launch_queries;
    b->hbm_budget = hbm_budget_bytes;
    b->n = srv->params.ring_dim;
    b->nb = 2 * (size_t)srv->params.record_bytes;
    auto drop_streams = [b] {
        for (hipStream_t st : {b->slots[0].compute, b->slots[1].compute, b->copy_in, b->copy_out})
            if (st) (void)hipStreamDestroy(st);
    };
    for (hipStream_t *st : {&b->slots[0].compute, &b->slots[1].compute, &b->copy_in, &b->copy_out}) {
        hipError_t e = hipStreamCreateWithFlags(st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            sm_err(err, err_cap, "HIP failure creating the batcher streams: %s", hipGetErrorString(e));
            drop_streams();
            delete b;
            return INSPIRE_ERR_DEVICE;
        }
    }
    for (Slot &s : b->slots) {
        int rc = alloc_slot(b, s, err, err_cap);
        if (rc) {
            drop_streams();
            delete b;
            return rc;
        }
    }
    b->worker = std::thread([b] { b->run(); });
    *out = b;
    return INSPIRE_OK;
}

void inspire_sm_batcher_destroy(inspire_sm_batcher *b) {
    if (!b) return;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->stopping = true;
    }
    b->cv_work.notify_all();
    b->worker.join();
    for (hipStream_t st : {b->copy_in, b->slots[0].compute, b->slots[1].compute, b->copy_out}) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete b;
}

int inspire_sm_batcher_register_key(inspire_sm_batcher *b, const uint8_t mask_seed[32], const uint32_t *key_q,
                                    const uint32_t *key_p, uint64_t *handle_out, char *err, size_t err_cap) {
    if (!b || !handle_out) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    const size_t bytes = 4 * b->n * 4;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        const size_t headroom = (size_t)256 << 20;
        if ((b->hbm_budget && b->resident + bytes > b->hbm_budget) || free_b < bytes + headroom) {
            sm_err(err, err_cap,
                   "key admission refused: seating %zu more bytes would exceed the device budget (%zu of %zu "
                   "bytes resident, %zu bytes free on the device)",
                   bytes, b->resident, b->hbm_budget, free_b);
            return INSPIRE_ERR_INVALID;
        }
    }
    // the seat's kernel is enqueued and not waited for: the launch that first carries one of its
    // requests waits for it on the device (launch())
    inspire_sm_seat *seat = nullptr;
    int rc = seat_create_enqueue(b->srv, mask_seed, key_q, key_p, &seat, err, err_cap);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(b->mu);
    const uint64_t h = b->next_handle++;
    b->seats[h] = std::shared_ptr<inspire_sm_seat>(seat, [](inspire_sm_seat *p) { inspire_sm_seat_destroy(p); });
    b->resident += bytes;
    *handle_out = h;
    return INSPIRE_OK;
}

int inspire_sm_batcher_unregister_key(inspire_sm_batcher *b, uint64_t handle, char *err, size_t err_cap) {
    if (!b) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    // queued requests keep the seat alive until they complete; otherwise it is destroyed when
    // `dropped` goes, after the lock (the destroy waits for the launches enqueued so far)
    std::shared_ptr<inspire_sm_seat> dropped;
    std::lock_guard<std::mutex> lk(b->mu);
    auto it = b->seats.find(handle);
    if (it == b->seats.end()) {
        sm_err(err, err_cap, "no key seated under handle %llu", (unsigned long long)handle);
        return INSPIRE_ERR_INVALID;
    }
    b->resident -= 4 * b->n * 4;
    dropped = std::move(it->second);
    b->seats.erase(it);
    return INSPIRE_OK;
}

int inspire_sm_batcher_submit(inspire_sm_batcher *b, uint64_t handle, const inspire_sm_table *const *tables,
                              const uint8_t *seeds, const uint32_t *bodies, size_t count, uint64_t *ticket_out,
                              char *err, size_t err_cap) {
    if (!b || !ticket_out || (count && (!tables || !seeds || !bodies))) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0 || count > b->max_launch) {
        sm_err(err, err_cap, "a request carries 1..=%zu queries, got %zu", b->max_launch, count);
        return INSPIRE_ERR_INVALID;
    }
    for (size_t k = 0; k < count; k++) {
        if (!tables[k]) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
        if (tables[k]->srv != b->srv) {
            sm_err(err, err_cap, "table or key seat belongs to another parameter set");
            return INSPIRE_ERR_INVALID;
        }
    }
    for (size_t k = 0; k < count * b->n; k++)
        if (bodies[k] >= b->srv->params.q) { sm_err(err, err_cap, "query does not fit the set"); return INSPIRE_ERR_INVALID; }
    Request r;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        auto it = b->seats.find(handle);
        if (it == b->seats.end()) {
            sm_err(err, err_cap, "no key seated under handle %llu", (unsigned long long)handle);
            return INSPIRE_ERR_INVALID;
        }
        r.seat = it->second;
        r.ticket = b->next_ticket++;
    }
    r.tables.assign(tables, tables + count);
    r.seeds.assign(seeds, seeds + count * 32);
    r.bodies.assign(bodies, bodies + count * b->n);
    r.submitted = clock_type::now();
    const uint64_t ticket = r.ticket;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->results[ticket] = Result();
        b->pending.push_back(std::move(r));
    }
    b->cv_work.notify_one();
    *ticket_out = ticket;
    return INSPIRE_OK;
}

int inspire_sm_batcher_wait(inspire_sm_batcher *b, uint64_t ticket, uint32_t *mask_out, uint32_t *body_out,
                            double *latency_ms, char *err, size_t err_cap) {
    if (!b) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    std::unique_lock<std::mutex> lk(b->mu);
    if (b->results.find(ticket) == b->results.end()) {
        sm_err(err, err_cap, "unknown ticket %llu", (unsigned long long)ticket);
        return INSPIRE_ERR_INVALID;
    }
    // found again on every wake-up: another waiter on the same ticket may have collected it
    auto it = b->results.end();
    b->cv_done.wait(lk, [&] {
        it = b->results.find(ticket);
        return it == b->results.end() || it->second.done;
    });
    if (it == b->results.end()) {
        sm_err(err, err_cap, "ticket %llu was collected by another wait", (unsigned long long)ticket);
        return INSPIRE_ERR_INVALID;
    }
    Result res = std::move(it->second);
    b->results.erase(it);
    lk.unlock();
    if (latency_ms) *latency_ms = res.latency_ms;
    if (res.rc != INSPIRE_OK) {
        sm_err(err, err_cap, "%s", res.err.c_str());
        return res.rc;
    }
    if (mask_out) memcpy(mask_out, res.mask.data(), res.mask.size() * 4);
    if (body_out) memcpy(body_out, res.body.data(), res.body.size() * 4);
    return INSPIRE_OK;
}

int inspire_sm_batcher_stats(inspire_sm_batcher *b, uint64_t *launches, uint64_t *queries, uint64_t *resident_bytes) {
    if (!b) return INSPIRE_ERR_INVALID;
    std::lock_guard<std::mutex> lk(b->mu);
    if (launches) *launches = b->launches.load();
    if (queries) *queries = b->queries.load();
    if (resident_bytes) *resident_bytes = b->resident;
    return INSPIRE_OK;
}

}  // extern "C"
