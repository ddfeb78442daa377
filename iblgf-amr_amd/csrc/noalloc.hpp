// Fixed device scratch of the allocation-free solve (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace iblgf_amd {

// Bytes of the scratch array each GPU holds (64 MiB; IBLGF_NOALLOC_SCRATCH_BYTES
// in include/iblgf_gpu.h).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;

// Looks up the device address of `device`'s scratch (that device must be
// current) once and keeps it; later calls return the kept address.
void* noalloc_scratch_init(int device);
// The kept address, or nullptr before noalloc_scratch_init(device).
void* noalloc_scratch(int device);
// Zeroes the first `bytes` (a multiple of 16) of a scratch region, by kernel.
void noalloc_clear(hipStream_t s, void* p, size_t bytes);

}  // namespace iblgf_amd
