// Kernel trace of the next launches: events recorded on each launch's own stream around its
// kernels, read after the caller's timed region. Recording waits for nothing, so a serving
// or benchmark loop keeps its pipelining while it is traced.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace inspire {

struct KernelTrace {
    int marks = 0;               // events per launch (kernels + 1)
    size_t cap = 0, n = 0;       // launches armed, launches recorded
    std::vector<hipEvent_t> ev;  // marks x cap

    ~KernelTrace() { clear(); }
    void clear() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        cap = n = 0;
    }
    // arm for `launches` launches of `kernels` kernels each; 0 launches disarms
    hipError_t arm(size_t launches, int kernels) {
        clear();
        marks = kernels + 1;
        ev.assign(launches * marks, nullptr);
        for (hipEvent_t &e : ev) {
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                e = nullptr;
                clear();
                return r;
            }
        }
        cap = launches;
        return hipSuccess;
    }
    // the event slots of the next launch, or null once the armed launches are recorded
    hipEvent_t *next() { return n < cap ? &ev[n * marks] : nullptr; }
    void done() { n++; }
    // durations of kernel k of each recorded launch (ms), once every recorded launch has
    // completed (the launches may sit on different streams, so each one's last event is
    // waited for, not only the last launch's); disarms
    hipError_t read(float *const *out, int kernels, size_t out_cap, size_t *count) {
        const size_t m = n < out_cap ? n : out_cap;
        hipError_t r = hipSuccess;
        for (size_t i = 0; i < n && r == hipSuccess; i++) r = hipEventSynchronize(ev[i * marks + marks - 1]);
        for (size_t i = 0; i < m && r == hipSuccess; i++)
            for (int k = 0; k < kernels && r == hipSuccess; k++)
                r = hipEventElapsedTime(&out[k][i], ev[i * marks + k], ev[i * marks + k + 1]);
        if (count) *count = r == hipSuccess ? m : 0;
        clear();
        return r;
    }
};

}  // namespace inspire
