// Device memory for library-allocated outputs (C's arrays, packed batches).
//
// Blocks come from a per-device cache of blocks earlier released with
// ops_free / ops_csr_free / ops_csr64_free; only a miss calls hipMalloc.
// New VRAM is cleared by the driver when it is mapped, so a fresh
// allocation of tens of GB costs as much as the product itself; repeated
// calls of similar size reuse their blocks instead.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace opsamd {

// Allocates at least `bytes` bytes on the current device.  On
// hipErrorOutOfMemory the device's cached blocks are released and the
// allocation is retried once.
hipError_t out_malloc(void **p, size_t bytes);

// Releases a block: library-allocated blocks go back to the cache (after
// the device is synchronised, as hipFree would), any other pointer is
// passed to hipFree.
hipError_t out_free(void *p);

}  // namespace opsamd
