// Workgroup placement on the MI355X's 8 XCDs (each with its own L2): consecutive workgroups of
// a launch are dispatched to the XCDs in turn, so workgroup b runs on XCD b mod 8.
#pragma once
#include <hip/hip_runtime.h>

namespace inspire {

constexpr int kXcds = 8;

// Position in a sorted order of the query that workgroup b of `count` takes: XCD x (= b mod 8)
// gets the x-th contiguous slice of the order, as many positions as it gets workgroups, in
// dispatch order (a bijection of [0, count)). Queries that share inputs (a shard, a seat) and
// sit next to each other in the order then run on one XCD at about the same time.
__host__ __device__ inline int xcd_slot(int b, int count) {
    const int x = b % kXcds, j = b / kXcds, m = count / kXcds, r = count % kXcds;
    return x * m + (x < r ? x : r) + j;
}

}  // namespace inspire
