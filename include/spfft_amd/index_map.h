/* Inspection hook of the MI355X build. No counterpart in the reference ABI:
 * it exposes the library's own restatement of convert_index_triplets
 * (reference src/compression/indices.hpp:120-186, parameters.cpp:173-179)
 * so the index maps can be compared bit for bit without a GPU.
 *
 * valueSlot: numValues ints, stick * dimZ + zStorage per value.
 * stickKey : up to numValues ints, xStorage * dimY + yStorage, ascending.
 * Returns SPFFT_SUCCESS, SPFFT_INVALID_PARAMETER_ERROR or
 * SPFFT_INVALID_INDICES_ERROR (as an int). */
#ifndef SPFFT_AMD_INDEX_MAP_H
#define SPFFT_AMD_INDEX_MAP_H

#include "spfft/config.h"

#ifdef __cplusplus
extern "C" {
#endif

SPFFT_EXPORT int spfft_amd_convert_index_triplets(int r2c, int dimX, int dimY, int dimZ,
                                                  int numValues, const int* triplets,
                                                  int* valueSlot, int* stickKey, int* numSticks,
                                                  int* zeroZeroStick);

#ifdef __cplusplus
}
#endif

#endif
