/* Feature switches of this build (reference: include/spfft/config.h.in:36-41).
 * ROCm GPU backend and single precision are always on; MPI is never compiled
 * in, so every *_distributed / *_communicator entry point is absent, exactly as
 * in a reference build configured with SPFFT_MPI=OFF. */
#ifndef SPFFT_CONFIG_H
#define SPFFT_CONFIG_H

#define SPFFT_ROCM
#define SPFFT_SINGLE_PRECISION

#include "spfft/spfft_export.h"

#endif
