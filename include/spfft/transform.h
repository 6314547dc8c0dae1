/* C interface of the transform object (double precision).
 * Signatures follow include/spfft/transform.h:68-332 of the reference; the
 * MPI-only entry points (transform.h:98-127, 334-342 there) are not part of
 * this build.
 *
 * Frequency-domain data: interleaved (re, im) pairs, one per index triplet, in
 * the order of the triplets given at creation.
 * Space-domain data: element (x, y, z) at (z * dimY + y) * dimX + x; complex
 * interleaved for SPFFT_TRANS_C2C, real for SPFFT_TRANS_R2C.
 * Backward: exp(+2 pi i k.n / N), unnormalised. Forward: exp(-2 pi i k.n / N),
 * scaled by 1 / (dimX dimY dimZ) only with SPFFT_FULL_SCALING.
 * With SPFFT_PU_GPU, input and output pointers may be device or host memory;
 * the pointer's attributes decide. */
#ifndef SPFFT_TRANSFORM_H
#define SPFFT_TRANSFORM_H

#include "spfft/config.h"
#include "spfft/errors.h"
#include "spfft/grid.h"
#include "spfft/types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void* SpfftTransform;

/* Create a transform that uses the buffers of `grid`. localZLength must equal
 * dimZ for a local grid. indices: 3 * numLocalElements ints (x,y,z), each in
 * [0, n-1] or, centred, in [n/2 - n + 1, n/2]; for R2C x is in [0, dimX/2]. */
SPFFT_EXPORT SpfftError spfft_transform_create(SpfftTransform* transform, SpfftGrid grid,
                                               SpfftProcessingUnitType processingUnit,
                                               SpfftTransformType transformType, int dimX, int dimY,
                                               int dimZ, int localZLength, int numLocalElements,
                                               SpfftIndexFormatType indexFormat,
                                               const int* indices);

/* Create a transform with a private grid sized for exactly this transform. */
SPFFT_EXPORT SpfftError spfft_transform_create_independent(
    SpfftTransform* transform, int maxNumThreads, SpfftProcessingUnitType processingUnit,
    SpfftTransformType transformType, int dimX, int dimY, int dimZ, int numLocalElements,
    SpfftIndexFormatType indexFormat, const int* indices);

SPFFT_EXPORT SpfftError spfft_transform_destroy(SpfftTransform transform);

/* Independent copy with freshly allocated buffers (deep-copies the grid). */
SPFFT_EXPORT SpfftError spfft_transform_clone(SpfftTransform transform,
                                              SpfftTransform* newTransform);

/* Forward transform of the internal space-domain buffer at inputLocation. */
SPFFT_EXPORT SpfftError spfft_transform_forward(SpfftTransform transform,
                                                SpfftProcessingUnitType inputLocation,
                                                double* output, SpfftScalingType scaling);

SPFFT_EXPORT SpfftError spfft_transform_forward_ptr(SpfftTransform transform, const double* input,
                                                    double* output, SpfftScalingType scaling);

/* Backward transform into the internal space-domain buffer at outputLocation. */
SPFFT_EXPORT SpfftError spfft_transform_backward(SpfftTransform transform, const double* input,
                                                 SpfftProcessingUnitType outputLocation);

SPFFT_EXPORT SpfftError spfft_transform_backward_ptr(SpfftTransform transform, const double* input,
                                                     double* output);

/* Pointer to the internal space-domain buffer on the host or the GPU. */
SPFFT_EXPORT SpfftError spfft_transform_get_space_domain(SpfftTransform transform,
                                                         SpfftProcessingUnitType dataLocation,
                                                         double** data);

SPFFT_EXPORT SpfftError spfft_transform_dim_x(SpfftTransform transform, int* dimX);
SPFFT_EXPORT SpfftError spfft_transform_dim_y(SpfftTransform transform, int* dimY);
SPFFT_EXPORT SpfftError spfft_transform_dim_z(SpfftTransform transform, int* dimZ);
SPFFT_EXPORT SpfftError spfft_transform_local_z_length(SpfftTransform transform, int* localZLength);
SPFFT_EXPORT SpfftError spfft_transform_local_slice_size(SpfftTransform transform, int* size);
SPFFT_EXPORT SpfftError spfft_transform_local_z_offset(SpfftTransform transform, int* offset);
SPFFT_EXPORT SpfftError spfft_transform_global_size(SpfftTransform transform,
                                                    long long int* globalSize);
/* Kept as in the reference (src/spfft/transform.cpp:433-445): reports the
 * local z length, not the element count. */
SPFFT_EXPORT SpfftError spfft_transform_num_local_elements(SpfftTransform transform,
                                                           int* numLocalElements);
SPFFT_EXPORT SpfftError spfft_transform_num_global_elements(SpfftTransform transform,
                                                            long long int* numGlobalElements);
SPFFT_EXPORT SpfftError spfft_transform_device_id(SpfftTransform transform, int* deviceId);
SPFFT_EXPORT SpfftError spfft_transform_num_threads(SpfftTransform transform, int* numThreads);
SPFFT_EXPORT SpfftError spfft_transform_execution_mode(SpfftTransform transform,
                                                       SpfftExecType* mode);
SPFFT_EXPORT SpfftError spfft_transform_set_execution_mode(SpfftTransform transform,
                                                           SpfftExecType mode);

#ifdef __cplusplus
}
#endif

#endif
