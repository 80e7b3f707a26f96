// Internal runtime objects of the special-modulus set, shared by the C ABI (sm_capi.cpp)
// and the continuous-batching queue (sm_batcher.cpp). Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/inspire_gpu.h"
#include "fence.hpp"
#include "sm.hpp"
#include "trace.hpp"
#include "xcd.hpp"

namespace inspire_sm {

inline void sm_err(char *err, size_t cap, const char *fmt, ...) {
    if (!err || cap == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, cap, fmt, ap);
    va_end(ap);
}

#define SM_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            inspire_sm::sm_err(err, err_cap, "HIP failure in %s: %s", #expr, hipGetErrorString(e_)); \
            return INSPIRE_ERR_DEVICE;                                                     \
        }                                                                                  \
    } while (0)

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = b;
        return hipMalloc(&p, b ? b : 8);
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Lays out a launch's jobs (built in batch order, qi = position) so that the queries of one
// seat, then of one table, run next to each other on one XCD: sorted by (seat, table, qi), and
// workgroup b takes sorted position xcd_slot(b). The chain re-reads the seat's four rows every
// step; with the queries of a seat together on an XCD those reads come from that XCD's L2.
inline void sm_order_jobs(inspire::SmJob *jobs, size_t count) {
    if (count < 2) return;
    std::vector<inspire::SmJob> sorted(jobs, jobs + count);
    std::sort(sorted.begin(), sorted.end(), [](const inspire::SmJob &a, const inspire::SmJob &b) {
        if (a.seat != b.seat) return a.seat < b.seat;
        if (a.step1 != b.step1) return a.step1 < b.step1;
        return a.qi < b.qi;
    });
    for (size_t b = 0; b < count; b++) jobs[b] = sorted[inspire::xcd_slot((int)b, (int)count)];
}

}  // namespace inspire_sm

struct inspire_sm_server {
    std::mutex mu;
    inspire_sm_params params{};
    inspire::SmCtxDev dev{};
    inspire_sm::Buf tw, cols;
    // staging of the host-buffer entry points
    inspire_sm::Buf jobs, seeds, bodies, mask, body;
    size_t cap = 0;
    // Job lists of the device entry points: a ring, so that launches on several streams can be
    // in flight at once. A list is written and read on one stream, reused by the next launch on
    // that stream while the same handles repeat, and rewritten only behind the last launch that
    // read it (its event, waited for on the device by the stream of the launch that rewrites it).
    struct JobList {
        inspire_sm::Buf buf;
        size_t cap = 0;
        hipEvent_t used = nullptr;  // recorded behind the last launch that read buf
        hipStream_t stream = nullptr;  // the stream buf was written and read on
        std::vector<inspire::SmJob> host;
    };
    static constexpr int kJobLists = 4;
    JobList djobs[kJobLists];
    int djobs_next = 0, djobs_last = -1;
    // staged response bytes of the host wire entry point
    inspire_sm::Buf wire;
    size_t wire_cap = 0;
    // the host-buffer entry points launch here and wait for this stream alone
    hipStream_t stream = nullptr;
    // kernel events of the next device-buffer launches, when armed
    inspire::KernelTrace trace;
    // an event behind every launch that reads this server's tables, seats or job list, recorded
    // under mu: freeing one of them waits for those launches only, not the whole device
    inspire::LaunchFence fence;
    // seats' device memory (4 x n words each), reused across registrations; a seat being
    // created stages its key limbs and mask seed in seat_in (2 n words + 32 bytes), on stream
    inspire::BlockPool seat_pool;
    inspire_sm::Buf seat_in;
    ~inspire_sm_server() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        (void)inspire::wait_for(fence.snapshot());
        for (JobList &j : djobs)
            if (j.used) (void)hipEventDestroy(j.used);
    }
    std::vector<hipEvent_t> launches_so_far() {
        std::lock_guard<std::mutex> lk(mu);
        return fence.snapshot();
    }
};

struct inspire_sm_table {
    inspire_sm_server *srv = nullptr;
    inspire_sm::Buf step1, cells;
    uint32_t rows = 0;
};

struct inspire_sm_seat {
    inspire_sm_server *srv = nullptr;
    inspire_sm::Buf words;
    // recorded on the server's stream behind the kernel that fills words; a launch on another
    // stream waits for it until it has been seen complete (settled)
    hipEvent_t ready = nullptr;
    bool settled = false;
    ~inspire_sm_seat() {
        if (ready) (void)hipEventDestroy(ready);
    }
};

namespace inspire_sm {
// inspire_sm_seat_create without the wait: the seat's kernel is enqueued on the server's stream
// and seat->ready recorded behind it (sm_capi.cpp).
int seat_create_enqueue(inspire_sm_server *s, const uint8_t mask_seed[32], const uint32_t *key_q,
                        const uint32_t *key_p, inspire_sm_seat **out, char *err, size_t err_cap);
}  // namespace inspire_sm
