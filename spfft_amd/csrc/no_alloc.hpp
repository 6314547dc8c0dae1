// Fixed device scratch of the allocation-free batched entry point
// (spfft_amd_multi_transform_noalloc): the batch's argument table and z block
// map live in a file-scope device array instead of memory allocated per
// batch, so the call makes no allocation of its own.
#pragma once

#include <cstddef>
#include <mutex>

namespace spfft {

// Device array per GPU: one argument table (224 bytes per FP64 transform,
// padded to 256 bytes in total) and one z block map (8 bytes per block of
// sticks).
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
// Transforms per call.
constexpr int kNoAllocMaxTransforms = 1024;
// Host image of the tables plus the device views of the user pointers.
constexpr size_t kNoAllocStageBytes =
    kNoAllocScratchBytes + 2 * kNoAllocMaxTransforms * sizeof(void*);

struct NoAllocScratch {
  void* device = nullptr;  // this GPU's copy of the array (kNoAllocScratchBytes)
  void* stage = nullptr;   // host stage shared by all GPUs (kNoAllocStageBytes)
  std::mutex* stageLock = nullptr;
};

// The scratch of `device`, which must be the current device. The first call
// on a device resolves the array's address there (hipGetSymbolAddress);
// later calls return the kept address.
NoAllocScratch noalloc_scratch(int device);

}  // namespace spfft
