/* Batched execution of independent transforms (single precision, available with SPFFT_SINGLE_PRECISION).
 * Signatures follow include/spfft/multi_transform_float.h:51-105 of the reference.
 * The transforms must not share a grid (SPFFT_INVALID_PARAMETER_ERROR
 * otherwise); each array holds numTransforms entries. */
#ifndef SPFFT_MULTI_TRANSFORM_FLOAT_H
#define SPFFT_MULTI_TRANSFORM_FLOAT_H

#include "spfft/config.h"
#include "spfft/transform_float.h"
#include "spfft/types.h"

#ifdef SPFFT_SINGLE_PRECISION
#ifdef __cplusplus
extern "C" {
#endif

SPFFT_EXPORT SpfftError spfft_float_multi_transform_forward(int numTransforms, SpfftFloatTransform* transforms,
                                                      const SpfftProcessingUnitType* inputLocations,
                                                      float* const* outputPointers,
                                                      const SpfftScalingType* scalingTypes);

SPFFT_EXPORT SpfftError spfft_float_multi_transform_forward_ptr(int numTransforms,
                                                          SpfftFloatTransform* transforms,
                                                          const float* const* inputPointers,
                                                          float* const* outputPointers,
                                                          const SpfftScalingType* scalingTypes);

SPFFT_EXPORT SpfftError spfft_float_multi_transform_backward(
    int numTransforms, SpfftFloatTransform* transforms, const float* const* inputPointers,
    const SpfftProcessingUnitType* outputLocations);

SPFFT_EXPORT SpfftError spfft_float_multi_transform_backward_ptr(int numTransforms,
                                                           SpfftFloatTransform* transforms,
                                                           const float* const* inputPointers,
                                                           float* const* outputPointers);

#ifdef __cplusplus
}
#endif
#endif /* SPFFT_SINGLE_PRECISION */

#endif
