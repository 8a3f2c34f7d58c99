// Allocation-free broadphase (noalloc.hip) for the C ABI.
#pragma once
#include <cstdint>

namespace cdx {

// detectPossibleCollisions on device buffers: aabb_aos (2n real3, mins then
// maxes) is shifted in place, the sorted unique pairs go to pairs_out
// (pair_cap entries).  Returns the pair count; throws cdx::Error with
// ERR_CAPACITY when n exceeds noalloc_max_shapes() or the bin entries or
// pairs exceed what the scratch array or pairs_out hold.  Synchronous on
// `stream` (a hipStream_t; null = the default stream).
uint64_t noalloc_broadphase(int device, int real_kind, void* aabb_aos, uint64_t n, int gx, int gy, int gz,
                            int64_t* pairs_out, uint64_t pair_cap, void* stream);
uint64_t noalloc_scratch_bytes();
uint64_t noalloc_max_shapes();

}  // namespace cdx
