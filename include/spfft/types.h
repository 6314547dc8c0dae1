/* Enumerations of the public API. Names and integer values match
 * include/spfft/types.h:33-117 of the reference so that binaries built against
 * either header agree. */
#ifndef SPFFT_TYPES_H
#define SPFFT_TYPES_H

#include "spfft/config.h"

/* MPI exchange flavour. Accepted for source compatibility only: this build has
 * no distributed path. */
enum SpfftExchangeType {
  SPFFT_EXCH_DEFAULT,
  SPFFT_EXCH_BUFFERED,
  SPFFT_EXCH_BUFFERED_FLOAT,
  SPFFT_EXCH_COMPACT_BUFFERED,
  SPFFT_EXCH_COMPACT_BUFFERED_FLOAT,
  SPFFT_EXCH_UNBUFFERED
};

/* Where a transform executes or where data lives. Bit flags: a grid may be
 * prepared for SPFFT_PU_HOST | SPFFT_PU_GPU. */
enum SpfftProcessingUnitType { SPFFT_PU_HOST = 1, SPFFT_PU_GPU = 2 };

/* Layout of the sparse frequency indices: interleaved x,y,z triplets. */
enum SpfftIndexFormatType { SPFFT_INDEX_TRIPLETS };

/* Complex-to-complex or real-to-complex (hermitian half in x). */
enum SpfftTransformType { SPFFT_TRANS_C2C, SPFFT_TRANS_R2C };

/* Forward-transform normalisation: none, or 1/(Nx*Ny*Nz). */
enum SpfftScalingType { SPFFT_NO_SCALING, SPFFT_FULL_SCALING };

/* SYNCHRONOUS: the call returns after the GPU work finished.
 * ASYNCHRONOUS: the call returns after enqueueing; the default stream is made
 * to wait for the transform. */
enum SpfftExecType { SPFFT_EXEC_SYNCHRONOUS, SPFFT_EXEC_ASYNCHRONOUS };

#ifndef __cplusplus
typedef enum SpfftExchangeType SpfftExchangeType;
typedef enum SpfftProcessingUnitType SpfftProcessingUnitType;
typedef enum SpfftTransformType SpfftTransformType;
typedef enum SpfftIndexFormatType SpfftIndexFormatType;
typedef enum SpfftScalingType SpfftScalingType;
typedef enum SpfftExecType SpfftExecType;
#endif

#endif
