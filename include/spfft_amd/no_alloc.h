/* Allocation-free batched transform of the MI355X build. No counterpart in
 * the reference ABI: it is for callers that cannot allocate device memory
 * during a call (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).
 *
 * spfft_amd_multi_transform_noalloc runs numTransforms double-precision GPU
 * transforms as one fused launch chain, like spfft_multi_transform_backward_ptr
 * (forward = 0) or spfft_multi_transform_forward_ptr (forward = 1, with one
 * scaling type per transform; scalingTypes is not read for backward). It
 * keeps the batch's argument table and block map in a fixed device array of
 * SPFFT_AMD_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, and
 * makes no device or host allocation: the transforms (and their grids) must
 * already exist, and inputs and outputs are caller-owned device memory.
 *
 * Requirements, each rejected with SPFFT_INVALID_PARAMETER_ERROR:
 *  - 1 <= numTransforms <= SPFFT_AMD_NOALLOC_MAX_TRANSFORMS, no two transforms
 *    on one grid (as the multi_transform functions);
 *  - GPU transforms on one device, all of one shape and type, on the
 *    fixed-length kernels (Nx = Ny in {32, 64, 96, 128, 192}), with non-empty
 *    index sets;
 *  - every input and output pointer in device memory;
 *  - the tables fit in SPFFT_AMD_NOALLOC_SCRATCH_BYTES (1 MiB): 224 bytes per
 *    transform (the total padded to 256) plus 8 bytes per block of sticks,
 *    a block holding 8 sticks at Nz = 128, 16 at Nz = 64 and 192, 32 at
 *    Nz = 32 and 96. Examples: up to 199 transforms of 128^3 with a radius-40
 *    sphere (5,025 sticks), up to 178 of 192^3 with radius 60 (11,289 sticks).
 * A null handle in the array is SPFFT_INVALID_PARAMETER_ERROR, as in the
 * multi_transform functions.
 *
 * The work is enqueued on the default stream; it is awaited before returning
 * when any transform is in SPFFT_EXEC_SYNCHRONOUS mode. All calls on one GPU
 * share its array, which keeps the last call's tables: calls whose GPU work
 * overlaps on one GPU (on other streams, or replays of graphs captured from
 * such calls running at the same time) race. Host threads are serialised
 * while the tables are written. */
#ifndef SPFFT_AMD_NO_ALLOC_H
#define SPFFT_AMD_NO_ALLOC_H

#include "spfft/config.h"
#include "spfft/errors.h"
#include "spfft/transform.h"
#include "spfft/types.h"

#define SPFFT_AMD_NOALLOC_SCRATCH_BYTES (1u << 20)
#define SPFFT_AMD_NOALLOC_MAX_TRANSFORMS 1024

#ifdef __cplusplus
extern "C" {
#endif

SPFFT_EXPORT SpfftError spfft_amd_multi_transform_noalloc(int numTransforms,
                                                          SpfftTransform* transforms,
                                                          const double* const* inputPointers,
                                                          double* const* outputPointers,
                                                          int forward,
                                                          const SpfftScalingType* scalingTypes);

#ifdef __cplusplus
}
#endif

#endif
