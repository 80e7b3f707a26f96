// Launch fence of one context or server: an event recorded on the stream of every launch that
// reads the owner's device objects (shard table, shards, seats, tables, job lists), so that
// freeing or rewriting such an object waits for exactly the launches enqueued before it
// instead of the whole device (other contexts' and other servers' work keeps running).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace inspire {

struct LaunchFence {
    std::vector<hipEvent_t> live;   // recorded after launches that may still run
    std::vector<hipEvent_t> spare;  // completed, ready to record again

    ~LaunchFence() {
        for (hipEvent_t e : live) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    }
    // Moves completed events back to the spare pool. Caller holds the owner's lock.
    void prune() {
        size_t k = 0;
        for (hipEvent_t e : live) {
            if (hipEventQuery(e) == hipSuccess)
                spare.push_back(e);
            else
                live[k++] = e;
        }
        live.resize(k);
    }
    // Records an event behind the launch just enqueued on s. Caller holds the owner's lock,
    // from the moment the launch read the owner's objects until this returns.
    hipError_t note(hipStream_t s) {
        prune();
        hipEvent_t e = nullptr;
        if (!spare.empty()) {
            e = spare.back();
            spare.pop_back();
        } else {
            hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
            if (r != hipSuccess) return r;
        }
        hipError_t r = hipEventRecord(e, s);
        if (r != hipSuccess) {
            spare.push_back(e);
            return r;
        }
        live.push_back(e);
        return hipSuccess;
    }
    // The launches enqueued so far (caller holds the owner's lock); wait on them with
    // wait_for() after releasing it. An event that completes and is recorded again by a later
    // launch meanwhile only makes the wait longer, never shorter.
    std::vector<hipEvent_t> snapshot() {
        prune();
        return live;
    }
};

// Device blocks of one size kept for reuse. hipFree waits for the whole device, so the seats
// that sessions open and close while launches run take their memory from here and give it back
// (after the fence wait of their destroy) instead of allocating and freeing it. Caller holds
// the owner's lock.
struct BlockPool {
    size_t block = 0;                 // bytes of every block
    std::vector<void *> free_blocks;  // no launch reads these any more
    ~BlockPool() {
        for (void *p : free_blocks) (void)hipFree(p);
    }
    hipError_t take(void **out) {
        if (!free_blocks.empty()) {
            *out = free_blocks.back();
            free_blocks.pop_back();
            return hipSuccess;
        }
        return hipMalloc(out, block);
    }
    void give(void *p) {
        if (p) free_blocks.push_back(p);
    }
};

inline hipError_t wait_for(const std::vector<hipEvent_t> &events) {
    hipError_t first = hipSuccess;
    for (hipEvent_t e : events) {
        hipError_t r = hipEventSynchronize(e);
        if (r != hipSuccess && first == hipSuccess) first = r;
    }
    return first;
}

}  // namespace inspire
