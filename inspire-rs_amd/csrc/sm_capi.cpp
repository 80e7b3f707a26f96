// Host runtime of the special-modulus set behind include/inspire_gpu.h: parameter
// checks (Params::validate, src/special_modulus/mod.rs:100-165), HBM-resident tables
// and seats, and the respond launches (server.rs:145-279).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "sm_runtime.hpp"

using namespace inspire;
using namespace inspire_sm;

namespace {

uint64_t pw(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    while (e) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return r;
}

bool is_prime(uint64_t m) {
    if (m < 2) return false;
    for (uint64_t d = 2; d * d <= m; d++)
        if (m % d == 0) return false;
    return true;
}

Pair32 pair(uint64_t w, uint64_t m) { return Pair32{(uint32_t)w, (uint32_t)((w << 32) / m)}; }

// -m^-1 mod 2^32 for odd m, by Newton's iteration (each step doubles the correct bits)
uint32_t neg_inverse(uint32_t m) {
    uint32_t inv = 1;
    for (int i = 0; i < 5; i++) inv *= 2u - m * inv;
    return 0u - inv;
}

int bitrev(int x, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// The root the reference picks (arith.rs:128-133): the first g whose (m-1)/2n-th power
// has order exactly 2n. Slot order, and so every transform-domain word, depends on it.
uint64_t first_root(uint64_t m, uint64_t n) {
    for (uint64_t g = 2; g < m; g++) {
        uint64_t c = pw(g, (m - 1) / (2 * n), m);
        if (pw(c, n, m) == m - 1) return c;
    }
    return 0;
}

}  // namespace

int inspire_sm::seat_create_enqueue(inspire_sm_server *s, const uint8_t mask_seed[32], const uint32_t *key_q,
                                    const uint32_t *key_p, inspire_sm_seat **out, char *err, size_t err_cap) {
    if (!s || !out || !mask_seed || !key_q || !key_p) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    const size_t n = s->params.ring_dim;
    for (size_t k = 0; k < n; k++)
        if (key_q[k] >= s->params.q || key_p[k] >= s->params.special_prime) {
            sm_err(err, err_cap, "packing key coefficient is not a residue");
            return INSPIRE_ERR_INVALID;
        }
    inspire_sm_seat *st = new (std::nothrow) inspire_sm_seat();
    if (!st) { sm_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    st->srv = s;
    // enqueued on the server's stream under its lock, with memory from the seat pool and the
    // inputs staged in seat_in (the next seat overwrites them only behind this kernel on the
    // same stream): nothing here waits
    hipError_t e = hipEventCreateWithFlags(&st->ready, hipEventDisableTiming);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->seat_pool.block = 4 * n * 4;
        void *mem = nullptr;
        if (e == hipSuccess) e = s->seat_pool.take(&mem);
        if (e == hipSuccess) {
            st->words.p = mem;
            st->words.bytes = s->seat_pool.block;
        }
        if (e == hipSuccess && !s->seat_in.p) e = s->seat_in.alloc(2 * n * 4 + 32);
        uint32_t *kq = s->seat_in.as<uint32_t>(), *kp = kq + n;
        uint8_t *seed = reinterpret_cast<uint8_t *>(kp + n);
        if (e == hipSuccess) e = hipMemcpyAsync(kq, key_q, n * 4, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(kp, key_p, n * 4, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(seed, mask_seed, 32, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = sm_launch_seat(s->dev, seed, kq, kp, st->words.as<uint32_t>(), s->stream);
        if (e == hipSuccess) e = hipEventRecord(st->ready, s->stream);
    }
    if (e != hipSuccess) {
        sm_err(err, err_cap, "HIP failure seating the packing key: %s", hipGetErrorString(e));
        delete st;
        return INSPIRE_ERR_DEVICE;
    }
    *out = st;
    return INSPIRE_OK;
}

extern "C" {

int inspire_sm_params_validate(const inspire_sm_params *p, char *err, size_t err_cap) {
    if (!p) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    const uint64_t d = p->ring_dim, gamma = p->gamma;
    if (d == 0 || (d & (d - 1)) || d < 16 || d > (1u << 16)) {
        sm_err(err, err_cap, "ring dimension %llu is not a power of two in [16, 2^16]", (unsigned long long)d);
        return INSPIRE_ERR_INVALID;
    }
    if (gamma == 0 || (gamma & (gamma - 1)) || gamma < 2 || gamma > d / 2) {
        sm_err(err, err_cap, "packed width %llu is not a power of two in [2, d/2]", (unsigned long long)gamma);
        return INSPIRE_ERR_INVALID;
    }
    if ((uint64_t)p->cell_bytes * 2 != gamma) {
        sm_err(err, err_cap, "a %u-byte cell fills %llu nibbles, not the packed width %llu", p->cell_bytes,
               (unsigned long long)p->cell_bytes * 2, (unsigned long long)gamma);
        return INSPIRE_ERR_INVALID;
    }
    if (p->record_bytes == 0 || p->record_bytes > p->cell_bytes) {
        sm_err(err, err_cap, "record of %u bytes does not fit a %u-byte cell", p->record_bytes, p->cell_bytes);
        return INSPIRE_ERR_INVALID;
    }
    for (uint64_t m : {p->q, p->special_prime})
        if (m >= (1ULL << 31) || !is_prime(m) || (m - 1) % (2 * d) != 0) {
            sm_err(err, err_cap, "%llu is not a prime below 2^31 that is 1 mod %llu", (unsigned long long)m,
                   (unsigned long long)(2 * d));
            return INSPIRE_ERR_INVALID;
        }
    if (p->special_prime <= p->q) {
        sm_err(err, err_cap, "special prime %llu is not above the ciphertext prime %llu",
               (unsigned long long)p->special_prime, (unsigned long long)p->q);
        return INSPIRE_ERR_INVALID;
    }
    if (p->q / 16 < (1u << 16)) {
        sm_err(err, err_cap, "ciphertext prime %llu leaves no room for noise", (unsigned long long)p->q);
        return INSPIRE_ERR_INVALID;
    }
    if (p->key_mask_version != INSPIRE_SM_KEY_MASK_VERSION) {
        sm_err(err, err_cap, "key-mask version %u is not the per-key-seed version %u", p->key_mask_version,
               INSPIRE_SM_KEY_MASK_VERSION);
        return INSPIRE_ERR_INVALID;
    }
    if (!(std::isfinite(p->sigma) && p->sigma >= 1.0 && p->sigma <= 64.0)) {
        sm_err(err, err_cap, "standard deviation %g is outside [1, 64]", p->sigma);
        return INSPIRE_ERR_INVALID;
    }
    return INSPIRE_OK;
}

int inspire_sm_server_create(const inspire_sm_params *params, inspire_sm_server **out, char *err, size_t err_cap) {
    if (!out) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    int rc = inspire_sm_params_validate(params, err, err_cap);
    if (rc) return rc;
    const uint64_t n = params->ring_dim, gamma = params->gamma, q = params->q, p = params->special_prime;
    // what the gfx950 kernels run: n/8 threads per query with 8 slots each (rings up to
    // 8192 in LDS) and 4-byte words of nibbles
    if (n > 8192) {
        sm_err(err, err_cap, "ring dimension %llu is outside the [16, 8192] the gfx950 kernels run",
               (unsigned long long)n);
        return INSPIRE_ERR_INVALID;
    }
    if (gamma < 8) {
        sm_err(err, err_cap, "packed width %llu is below the 8 the gfx950 kernels run", (unsigned long long)gamma);
        return INSPIRE_ERR_INVALID;
    }
    int logn = 0;
    while ((1ULL << logn) < n) logn++;
    const uint64_t psi_q = first_root(q, n), psi_p = first_root(p, n);
    if (!psi_q || !psi_p) { sm_err(err, err_cap, "no primitive %llu-th root", (unsigned long long)(2 * n)); return INSPIRE_ERR_INVALID; }
    // twiddles: psi^(+-bitrev(k)) for both primes, with companions
    std::vector<Pair32> tw(4 * n);
    {
        const uint64_t ms[2] = {q, p}, ps[2] = {psi_q, psi_p};
        for (int l = 0; l < 2; l++) {
            const uint64_t m = ms[l], inv = pw(ps[l], 2 * n - 1, m);
            std::vector<uint64_t> fp(n), ip(n);
            uint64_t a = 1, b = 1;
            for (uint64_t k = 0; k < n; k++) {
                fp[k] = a;
                ip[k] = b;
                a = a * ps[l] % m;
                b = b * inv % m;
            }
            for (uint64_t k = 0; k < n; k++) {
                tw[(2 * l) * n + k] = pair(fp[bitrev((int)k, logn)], m);
                tw[(2 * l + 1) * n + k] = pair(ip[bitrev((int)k, logn)], m);
            }
        }
    }
    inspire_sm_server *s = new (std::nothrow) inspire_sm_server();
    if (!s) { sm_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    s->params = *params;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = s->tw.alloc(tw.size() * sizeof(Pair32));
    if (e == hipSuccess) e = hipMemcpy(s->tw.p, tw.data(), tw.size() * sizeof(Pair32), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = s->cols.alloc(gamma * n);
    if (e != hipSuccess) {
        sm_err(err, err_cap, "HIP failure creating the special-modulus server: %s", hipGetErrorString(e));
        delete s;
        return INSPIRE_ERR_DEVICE;
    }
    SmCtxDev &c = s->dev;
    c.logn = logn;
    c.n = (int)n;
    c.gamma = (int)gamma;
    c.nb = 2 * (int)params->record_bytes;
    c.cell_bytes = (int)params->cell_bytes;
    c.q = (uint32_t)q;
    c.p = (uint32_t)p;
    c.half_q = (uint32_t)(q / 2);
    c.half_p = (uint32_t)(p / 2);
    c.p_minus_q = (uint32_t)(p - q);
    c.q_barrett = (uint32_t)((1ULL << 32) / q);
    c.neg_off = (uint32_t)((q - p % q) % q);
    c.g = (uint32_t)(2 * n / gamma + 1);
    c.wide_q = q > kSmLazyQMax ? 1 : 0;  // every prime below 2^31 that Params::validate accepts
    c.q_recip = UINT64_MAX / q;
    c.p_recip = UINT64_MAX / p;
    c.zone_q = (UINT64_MAX / q) * q;
    c.zone_p = (UINT64_MAX / p) * p;
    const uint64_t inv_n_q = pw(n, q - 2, q), pinv = pw(p % q, q - 2, q);
    const uint64_t r_q = (1ULL << 32) % q, r_p = (1ULL << 32) % p;
    c.qninv = neg_inverse((uint32_t)q);
    c.pninv = neg_inverse((uint32_t)p);
    c.pinv_n = pair(pinv * inv_n_q % q, q);
    c.pinv = pair(pinv, q);
    c.table_scale = pair(inv_n_q * pw(gamma, q - 2, q) % q * r_q % q, q);
    c.seat_wq = pair(n % q * r_q % q, q);
    c.seat_kq = pair(r_q, q);
    c.seat_p = pair(pw(n, p - 2, p) * r_p % p, p);
    Pair32 *t = s->tw.as<Pair32>();
    c.twf_q = t;
    c.twi_q = t + n;
    c.twf_p = t + 2 * n;
    c.twi_p = t + 3 * n;
    *out = s;
    return INSPIRE_OK;
}

void inspire_sm_server_destroy(inspire_sm_server *s) {
    if (!s) return;
    (void)wait_for(s->launches_so_far());
    delete s;
}

int inspire_sm_table_create(inspire_sm_server *s, const uint8_t *cells, size_t len, inspire_sm_table **out, char *err,
                            size_t err_cap) {
    if (!s || !out || (len && !cells)) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
    const size_t n = s->params.ring_dim, cell = s->params.cell_bytes, rec = s->params.record_bytes;
    if (len % cell != 0 || len / cell > n) {
        sm_err(err, err_cap, "%zu bytes are not at most %zu cells of %zu bytes", len, n, cell);
        return INSPIRE_ERR_INVALID;
    }
    const size_t rows = len / cell;
    for (size_t m = 0; m < rows; m++)
        for (size_t b = rec; b < cell; b++)
            if (cells[m * cell + b] != 0) {
                sm_err(err, err_cap, "cell %zu has bytes past its %zu-byte record", m, rec);
                return INSPIRE_ERR_INVALID;
            }
    inspire_sm_table *t = new (std::nothrow) inspire_sm_table();
    if (!t) { sm_err(err, err_cap, "out of host memory"); return INSPIRE_ERR_DEVICE; }
    t->srv = s;
    t->rows = (uint32_t)rows;
    std::lock_guard<std::mutex> lk(s->mu);
    hipError_t e = t->cells.alloc(n * cell);
    if (e == hipSuccess) e = t->step1.alloc((size_t)s->params.gamma * n * 4);
    if (e == hipSuccess) e = hipMemset(t->cells.p, 0, n * cell);
    if (e == hipSuccess && len) e = hipMemcpy(t->cells.p, cells, len, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = sm_launch_table(s->dev, t->cells.as<uint8_t>(), s->cols.as<int8_t>(), t->step1.as<uint32_t>(), s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        sm_err(err, err_cap, "HIP failure building the step-1 table: %s", hipGetErrorString(e));
        delete t;
        return INSPIRE_ERR_DEVICE;
    }
    *out = t;
    return INSPIRE_OK;
}

uint32_t inspire_sm_table_rows(const inspire_sm_table *t) { return t ? t->rows : 0; }
size_t inspire_sm_table_resident_bytes(const inspire_sm_table *t) { return t ? t->step1.bytes + t->cells.bytes : 0; }

void inspire_sm_table_destroy(inspire_sm_table *t) {
    if (!t) return;
    (void)wait_for(t->srv->launches_so_far());
    delete t;
}

int inspire_sm_seat_create(inspire_sm_server *s, const uint8_t mask_seed[32], const uint32_t *key_q,
                           const uint32_t *key_p, inspire_sm_seat **out, char *err, size_t err_cap) {
    inspire_sm_seat *st = nullptr;
    int rc = seat_create_enqueue(s, mask_seed, key_q, key_p, &st, err, err_cap);
    if (rc) return rc;
    hipError_t e = hipEventSynchronize(st->ready);
    if (e != hipSuccess) {
        sm_err(err, err_cap, "HIP failure seating the packing key: %s", hipGetErrorString(e));
        inspire_sm_seat_destroy(st);
        return INSPIRE_ERR_DEVICE;
    }
    st->settled = true;
    *out = st;
    return INSPIRE_OK;
}

size_t inspire_sm_seat_resident_bytes(const inspire_sm_seat *st) { return st ? st->words.bytes : 0; }

void inspire_sm_seat_destroy(inspire_sm_seat *st) {
    if (!st) return;
    inspire_sm_server *s = st->srv;
    (void)wait_for(s->launches_so_far());
    {
        // no launch enqueued so far reads the seat any more: its memory goes back to the pool
        std::lock_guard<std::mutex> lk(s->mu);
        s->seat_pool.give(st->words.p);
        st->words.p = nullptr;
    }
    delete st;
}

}  // extern "C"

namespace {

int check_handles(inspire_sm_server *s, size_t count, const inspire_sm_table *const *tables,
                  const inspire_sm_seat *const *seats, std::vector<SmJob> &jobs, char *err, size_t err_cap) {
    jobs.resize(count);
    for (size_t k = 0; k < count; k++) {
        const inspire_sm_table *t = tables[k];
        const inspire_sm_seat *st = seats[k];
        if (!t || !st) { sm_err(err, err_cap, "null argument"); return INSPIRE_ERR_INVALID; }
        if (t->srv != s || st->srv != s) {
            sm_err(err, err_cap, "table or key seat belongs to another parameter set");
            return INSPIRE_ERR_INVALID;
        }
        jobs[k] = SmJob{t->step1.as<uint32_t>(), t->cells.as<uint8_t>(), st->words.as<uint32_t>(), (int)k};
    }
    sm_order_jobs(jobs.data(), count);
    return INSPIRE_OK;
}

// a device-buffer launch, with events around its kernel while a trace is armed (no wait)
hipError_t launch_traced(inspire_sm_server *s, const SmBatchDev &b, hipStream_t st) {
    hipEvent_t *tr = s->trace.next();
    if (!tr) return sm_launch_respond(s->dev, b, st);
    hipError_t e = hipEventRecord(tr[0], st);
    if (e == hipSuccess) e = sm_launch_respond(s->dev, b, st);
    if (e == hipSuccess) e = hipEventRecord(tr[1], st);
    if (e == hipSuccess) s->trace.done();
    return e;
}

}  // namespace

extern "C" {

int inspire_sm_respond_batch(inspire_sm_server *s, size_t count, const inspire_sm_table *const *tables,
                             const inspire_sm_seat *const *seats, const uint8_t *seeds, const uint32_t *bodies,
                             uint32_t *mask_out, uint32_t *body_out, char *err, size_t err_cap) {
    if (!s || (count && (!tables || !seats || !seeds || !bodies || !mask_out || !body_out))) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0) return INSPIRE_OK;
    if (count > INSPIRE_MAX_BATCH) {
        sm_err(err, err_cap, "a batch holds at most %u queries, got %zu", INSPIRE_MAX_BATCH, count);
        return INSPIRE_ERR_INVALID;
    }
    const size_t n = s->params.ring_dim, nb = 2 * (size_t)s->params.record_bytes;
    std::vector<SmJob> jobs;
    int rc = check_handles(s, count, tables, seats, jobs, err, err_cap);
    if (rc) return rc;
    for (size_t k = 0; k < count * n; k++)
        if (bodies[k] >= s->params.q) { sm_err(err, err_cap, "query does not fit the set"); return INSPIRE_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (count > s->cap) {
        SM_TRY(s->jobs.alloc(count * sizeof(SmJob)));
        SM_TRY(s->seeds.alloc(count * 32));
        SM_TRY(s->bodies.alloc(count * n * 4));
        SM_TRY(s->mask.alloc(count * n * 4));
        SM_TRY(s->body.alloc(count * nb * 4));
        s->cap = count;
    }
    SM_TRY(hipMemcpy(s->jobs.p, jobs.data(), count * sizeof(SmJob), hipMemcpyHostToDevice));
    SM_TRY(hipMemcpy(s->seeds.p, seeds, count * 32, hipMemcpyHostToDevice));
    SM_TRY(hipMemcpy(s->bodies.p, bodies, count * n * 4, hipMemcpyHostToDevice));
    SmBatchDev b{(int)count, s->jobs.as<SmJob>(), s->seeds.as<uint8_t>(), s->bodies.as<uint32_t>(),
                 s->mask.as<uint32_t>(), s->body.as<uint32_t>()};
    SM_TRY(sm_launch_respond(s->dev, b, s->stream));
    SM_TRY(hipStreamSynchronize(s->stream));
    SM_TRY(hipMemcpy(mask_out, s->mask.p, count * n * 4, hipMemcpyDeviceToHost));
    SM_TRY(hipMemcpy(body_out, s->body.p, count * nb * 4, hipMemcpyDeviceToHost));
    return INSPIRE_OK;
}

int inspire_sm_respond(inspire_sm_server *s, const inspire_sm_table *table, const inspire_sm_seat *seat,
                       const uint8_t seed[32], const uint32_t *body, uint32_t *mask_out, uint32_t *body_out,
                       char *err, size_t err_cap) {
    return inspire_sm_respond_batch(s, 1, &table, &seat, seed, body, mask_out, body_out, err, err_cap);
}

// The device entry points' job list, uploaded only when the handles change, so a repeated
// batch launches with nothing but the kernel on the stream. A launch of this server on any
// stream may still read the old list, so a change waits for the server's launches enqueued so
// far (its fence), not for the whole device. Caller holds s->mu.
// The device job list of a launch on stream st (caller holds s->mu): the newest list when it
// was written on st and the handles repeat it, else the next list of the ring, rewritten on st
// behind the last launch that read it (all on one stream, so the last one is the latest).
// *used receives the list's event, which the caller records behind its launch.
static int upload_device_jobs(inspire_sm_server *s, const std::vector<SmJob> &jobs, hipStream_t st,
                              const SmJob **dev_jobs, hipEvent_t *used, char *err, size_t err_cap) {
    if (s->djobs_last >= 0) {
        inspire_sm_server::JobList &l = s->djobs[s->djobs_last];
        if (l.stream == st && jobs.size() == l.host.size() &&
            std::memcmp(jobs.data(), l.host.data(), jobs.size() * sizeof(SmJob)) == 0) {
            // ordered behind the list's copy already (same stream); the wait only matters for a
            // stream created with a destroyed one's handle
            SM_TRY(hipStreamWaitEvent(st, l.used, 0));
            *dev_jobs = l.buf.as<SmJob>();
            *used = l.used;
            return INSPIRE_OK;
        }
    }
    const int k = s->djobs_next;
    inspire_sm_server::JobList &l = s->djobs[k];
    if (!l.used) SM_TRY(hipEventCreateWithFlags(&l.used, hipEventDisableTiming));
    if (jobs.size() > l.cap) {
        // a larger list: the old buffer may still be read by launches in flight
        SM_TRY(wait_for(s->fence.snapshot()));
        SM_TRY(l.buf.alloc(jobs.size() * sizeof(SmJob)));
        l.cap = jobs.size();
    } else {
        SM_TRY(hipStreamWaitEvent(st, l.used, 0));
    }
    SM_TRY(hipMemcpyAsync(l.buf.p, jobs.data(), jobs.size() * sizeof(SmJob), hipMemcpyHostToDevice, st));
    l.host = jobs;
    l.stream = st;
    s->djobs_last = k;
    s->djobs_next = (k + 1) % inspire_sm_server::kJobLists;
    *dev_jobs = l.buf.as<SmJob>();
    *used = l.used;
    return INSPIRE_OK;
}

size_t inspire_sm_response_wire_bytes(const inspire_sm_server *s) {
    if (!s) return 0;
    return sm_wire((uint32_t)s->params.ring_dim, 2 * (uint32_t)s->params.record_bytes, (uint32_t)s->params.q).total;
}

int inspire_sm_respond_batch_wire(inspire_sm_server *s, size_t count, const inspire_sm_table *const *tables,
                                  const inspire_sm_seat *const *seats, const uint8_t *seeds, const uint32_t *bodies,
                                  uint8_t *wire_out, size_t wire_stride, char *err, size_t err_cap) {
    if (!s || (count && (!tables || !seats || !seeds || !bodies || !wire_out))) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0) return INSPIRE_OK;
    if (count > INSPIRE_MAX_BATCH) {
        sm_err(err, err_cap, "a batch holds at most %u queries, got %zu", INSPIRE_MAX_BATCH, count);
        return INSPIRE_ERR_INVALID;
    }
    const size_t n = s->params.ring_dim;
    const SmWire wl = sm_wire((uint32_t)n, 2 * (uint32_t)s->params.record_bytes, (uint32_t)s->params.q);
    if (wire_stride < wl.total) {
        sm_err(err, err_cap, "response stride %zu is below the %u-byte response", wire_stride, wl.total);
        return INSPIRE_ERR_INVALID;
    }
    std::vector<SmJob> jobs;
    int rc = check_handles(s, count, tables, seats, jobs, err, err_cap);
    if (rc) return rc;
    for (size_t k = 0; k < count * n; k++)
        if (bodies[k] >= s->params.q) { sm_err(err, err_cap, "query does not fit the set"); return INSPIRE_ERR_INVALID; }
    const size_t dstride = (wl.total + 3) & ~(size_t)3;
    std::lock_guard<std::mutex> lk(s->mu);
    if (count > s->cap) {
        SM_TRY(s->jobs.alloc(count * sizeof(SmJob)));
        SM_TRY(s->seeds.alloc(count * 32));
        SM_TRY(s->bodies.alloc(count * n * 4));
        SM_TRY(s->mask.alloc(count * n * 4));
        SM_TRY(s->body.alloc(count * 2 * (size_t)s->params.record_bytes * 4));
        s->cap = count;
    }
    if (s->wire_cap < count * dstride) {
        SM_TRY(s->wire.alloc(count * dstride));
        s->wire_cap = count * dstride;
    }
    SM_TRY(hipMemcpy(s->jobs.p, jobs.data(), count * sizeof(SmJob), hipMemcpyHostToDevice));
    SM_TRY(hipMemcpy(s->seeds.p, seeds, count * 32, hipMemcpyHostToDevice));
    SM_TRY(hipMemcpy(s->bodies.p, bodies, count * n * 4, hipMemcpyHostToDevice));
    SmBatchDev b{(int)count, s->jobs.as<SmJob>(), s->seeds.as<uint8_t>(), s->bodies.as<uint32_t>(),
                 s->mask.as<uint32_t>(), nullptr};
    b.wire_out = s->wire.as<uint8_t>();
    b.wire_stride = dstride;
    SM_TRY(sm_launch_respond(s->dev, b, s->stream));
    SM_TRY(hipStreamSynchronize(s->stream));
    SM_TRY(hipMemcpy2D(wire_out, wire_stride, s->wire.p, dstride, wl.total, count, hipMemcpyDeviceToHost));
    return INSPIRE_OK;
}

int inspire_sm_respond_batch_device_wire(inspire_sm_server *s, size_t count, const inspire_sm_table *const *tables,
                                         const inspire_sm_seat *const *seats, const uint8_t *d_seeds,
                                         const uint32_t *d_bodies, uint32_t *d_scratch, uint8_t *d_wire_out,
                                         size_t wire_stride, void *stream, char *err, size_t err_cap) {
    if (!s || (count && (!tables || !seats || !d_seeds || !d_bodies || !d_scratch || !d_wire_out))) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0) return INSPIRE_OK;
    if (count > INSPIRE_MAX_BATCH) {
        sm_err(err, err_cap, "a batch holds at most %u queries, got %zu", INSPIRE_MAX_BATCH, count);
        return INSPIRE_ERR_INVALID;
    }
    const size_t n = s->params.ring_dim;
    const SmWire wl = sm_wire((uint32_t)n, 2 * (uint32_t)s->params.record_bytes, (uint32_t)s->params.q);
    if (wire_stride < ((wl.total + 3) & ~3u) || wire_stride % 4 != 0) {
        sm_err(err, err_cap, "response stride %zu must be a multiple of 4 holding %u bytes", wire_stride,
               (wl.total + 3) & ~3u);
        return INSPIRE_ERR_INVALID;
    }
    std::vector<SmJob> jobs;
    int rc = check_handles(s, count, tables, seats, jobs, err, err_cap);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(s->mu);
    const SmJob *dj = nullptr;
    hipEvent_t used = nullptr;
    rc = upload_device_jobs(s, jobs, st, &dj, &used, err, err_cap);
    if (rc) return rc;
    SmBatchDev b{(int)count, dj, d_seeds, d_bodies, d_scratch, nullptr};
    b.wire_out = d_wire_out;
    b.wire_stride = wire_stride;
    SM_TRY(launch_traced(s, b, st));
    SM_TRY(hipEventRecord(used, st));
    SM_TRY(s->fence.note(st));
    return INSPIRE_OK;
}

int inspire_sm_respond_batch_device(inspire_sm_server *s, size_t count, const inspire_sm_table *const *tables,
                                    const inspire_sm_seat *const *seats, const uint8_t *d_seeds,
                                    const uint32_t *d_bodies, uint32_t *d_mask_out, uint32_t *d_body_out,
                                    void *stream, float *kernel_ms, char *err, size_t err_cap) {
    if (!s || (count && (!tables || !seats || !d_seeds || !d_bodies || !d_mask_out || !d_body_out))) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    if (count == 0) return INSPIRE_OK;
    if (count > INSPIRE_MAX_BATCH) {
        sm_err(err, err_cap, "a batch holds at most %u queries, got %zu", INSPIRE_MAX_BATCH, count);
        return INSPIRE_ERR_INVALID;
    }
    std::vector<SmJob> jobs;
    int rc = check_handles(s, count, tables, seats, jobs, err, err_cap);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(s->mu);
    const SmJob *dj = nullptr;
    hipEvent_t used = nullptr;
    rc = upload_device_jobs(s, jobs, st, &dj, &used, err, err_cap);
    if (rc) return rc;
    SmBatchDev b{(int)count, dj, d_seeds, d_bodies, d_mask_out, d_body_out};
    if (!kernel_ms) {
        SM_TRY(launch_traced(s, b, st));
        SM_TRY(hipEventRecord(used, st));
        SM_TRY(s->fence.note(st));
        return INSPIRE_OK;
    }
    hipEvent_t ev[2];
    SM_TRY(hipEventCreate(&ev[0]));
    SM_TRY(hipEventCreate(&ev[1]));
    hipError_t e = hipEventRecord(ev[0], st);
    if (e == hipSuccess) e = sm_launch_respond(s->dev, b, st);
    if (e == hipSuccess) e = hipEventRecord(ev[1], st);
    if (e == hipSuccess) e = hipEventRecord(used, st);
    if (e == hipSuccess) e = s->fence.note(st);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(kernel_ms, ev[0], ev[1]);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    SM_TRY(e);
    return INSPIRE_OK;
}

int inspire_sm_server_trace(inspire_sm_server *s, size_t launches, char *err, size_t err_cap) {
    if (!s) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    if (launches > (1u << 20)) {
        sm_err(err, err_cap, "a trace holds at most %u launches, asked for %zu", 1u << 20, launches);
        return INSPIRE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    SM_TRY(s->trace.arm(launches, 1));
    return INSPIRE_OK;
}

int inspire_sm_server_kernel_times(inspire_sm_server *s, float *kernel_ms, size_t cap, size_t *count, char *err,
                                   size_t err_cap) {
    if (!s || !count || (cap && !kernel_ms)) {
        sm_err(err, err_cap, "null argument");
        return INSPIRE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    float *const out[1] = {kernel_ms};
    SM_TRY(s->trace.read(out, 1, cap, count));
    return INSPIRE_OK;
}

}  // extern "C"
