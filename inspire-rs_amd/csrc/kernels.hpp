// Launchers of the respond-path kernels (definitions in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace inspire {

hipError_t launch_respond(const CtxDev &c, const BatchDev &b, hipStream_t s);
// the three stages of launch_respond: shard grouping (grouped launches only), DB scan, packing
hipError_t launch_group_only(const CtxDev &c, const BatchDev &b, hipStream_t s);
hipError_t launch_scan_only(const CtxDev &c, const BatchDev &b, hipStream_t s);
hipError_t launch_chain_only(const CtxDev &c, const BatchDev &b, hipStream_t s);
hipError_t launch_expand_mask(int logn, int rows, const uint8_t *seed, uint64_t stream, int dropped,
                              const uint64_t *psi, uint64_t inv_n, Pair *out, hipStream_t s);
hipError_t launch_seat(int logn, int rows, const uint64_t *y, const uint64_t *psi, uint64_t inv_n, Pair *out,
                       hipStream_t s);
hipError_t launch_store_build(const CtxDev &c, const uint16_t *records, uint16_t *col_scratch, const uint64_t *tag_key,
                              Pair *store, uint64_t *tags, hipStream_t s);

hipError_t launch_fold_rows(Pair *rows, int n, uint64_t kx, uint64_t k0, uint64_t k1, hipStream_t s);

}  // namespace inspire
