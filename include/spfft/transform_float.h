/* C interface of the transform object (single precision, available with SPFFT_SINGLE_PRECISION).
 * Signatures follow include/spfft/transform_float.h:67-341 of the reference; the
 * MPI-only entry points are not part of
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
#ifndef SPFFT_TRANSFORM_FLOAT_H
#define SPFFT_TRANSFORM_FLOAT_H

#include "spfft/config.h"
#include "spfft/errors.h"
#include "spfft/grid_float.h"
#include "spfft/types.h"

#ifdef SPFFT_SINGLE_PRECISION
#ifdef __cplusplus
extern "C" {
#endif

typedef void* SpfftFloatTransform;

/* Create a transform that uses the buffers of `grid`. localZLength must equal
 * dimZ for a local grid. indices: 3 * numLocalElements ints (x,y,z), each in
 * [0, n-1] or, centred, in [n/2 - n + 1, n/2]; for R2C x is in [0, dimX/2]. */
SPFFT_EXPORT SpfftError spfft_float_transform_create(SpfftFloatTransform* transform, SpfftFloatGrid grid,
                                               SpfftProcessingUnitType processingUnit,
                                               SpfftTransformType transformType, int dimX, int dimY,
                                               int dimZ, int localZLength, int numLocalElements,
                                               SpfftIndexFormatType indexFormat,
                                               const int* indices);

/* Create a transform with a private grid sized for exactly this transform. */
SPFFT_EXPORT SpfftError spfft_float_transform_create_independent(
    SpfftFloatTransform* transform, int maxNumThreads, SpfftProcessingUnitType processingUnit,
    SpfftTransformType transformType, int dimX, int dimY, int dimZ, int numLocalElements,
    SpfftIndexFormatType indexFormat, const int* indices);

SPFFT_EXPORT SpfftError spfft_float_transform_destroy(SpfftFloatTransform transform);

/* Independent copy with freshly allocated buffers (deep-copies the grid). */
SPFFT_EXPORT SpfftError spfft_float_transform_clone(SpfftFloatTransform transform,
                                              SpfftFloatTransform* newTransform);

/* Forward transform of the internal space-domain buffer at inputLocation. */
SPFFT_EXPORT SpfftError spfft_float_transform_forward(SpfftFloatTransform transform,
                                                SpfftProcessingUnitType inputLocation,
                                                float* output, SpfftScalingType scaling);

SPFFT_EXPORT SpfftError spfft_float_transform_forward_ptr(SpfftFloatTransform transform, const float* input,
                                                    float* output, SpfftScalingType scaling);

/* Backward transform into the internal space-domain buffer at outputLocation. */
SPFFT_EXPORT SpfftError spfft_float_transform_backward(SpfftFloatTransform transform, const float* input,
                                                 SpfftProcessingUnitType outputLocation);

SPFFT_EXPORT SpfftError spfft_float_transform_backward_ptr(SpfftFloatTransform transform, const float* input,
                                                     float* output);

/* Pointer to the internal space-domain buffer on the host or the GPU. */
SPFFT_EXPORT SpfftError spfft_float_transform_get_space_domain(SpfftFloatTransform transform,
                                                         SpfftProcessingUnitType dataLocation,
                                                         float** data);

SPFFT_EXPORT SpfftError spfft_float_transform_dim_x(SpfftFloatTransform transform, int* dimX);
SPFFT_EXPORT SpfftError spfft_float_transform_dim_y(SpfftFloatTransform transform, int* dimY);
SPFFT_EXPORT SpfftError spfft_float_transform_dim_z(SpfftFloatTransform transform, int* dimZ);
SPFFT_EXPORT SpfftError spfft_float_transform_local_z_length(SpfftFloatTransform transform, int* localZLength);
SPFFT_EXPORT SpfftError spfft_float_transform_local_slice_size(SpfftFloatTransform transform, int* size);
SPFFT_EXPORT SpfftError spfft_float_transform_local_z_offset(SpfftFloatTransform transform, int* offset);
SPFFT_EXPORT SpfftError spfft_float_transform_global_size(SpfftFloatTransform transform,
                                                    long long int* globalSize);
/* Kept as in the reference (src/spfft/transform_float.cpp:433-446): reports the
 * local z length, not the element count. */
SPFFT_EXPORT SpfftError spfft_float_transform_num_local_elements(SpfftFloatTransform transform,
                                                           int* numLocalElements);
SPFFT_EXPORT SpfftError spfft_float_transform_num_global_elements(SpfftFloatTransform transform,
                                                            long long int* numGlobalElements);
SPFFT_EXPORT SpfftError spfft_float_transform_device_id(SpfftFloatTransform transform, int* deviceId);
SPFFT_EXPORT SpfftError spfft_float_transform_num_threads(SpfftFloatTransform transform, int* numThreads);
SPFFT_EXPORT SpfftError spfft_float_transform_execution_mode(SpfftFloatTransform transform,
                                                       SpfftExecType* mode);
SPFFT_EXPORT SpfftError spfft_float_transform_set_execution_mode(SpfftFloatTransform transform,
                                                           SpfftExecType mode);

#ifdef __cplusplus
}
#endif
#endif /* SPFFT_SINGLE_PRECISION */

#endif
