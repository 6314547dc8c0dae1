/* C interface of the grid object (single precision, available with SPFFT_SINGLE_PRECISION).
 * Signatures follow include/spfft/grid_float.h:60-178 of the reference; the
 * MPI-only constructor and communicator accessor are not part of this build.
 *
 * A grid owns the device work buffers sized for its maximum dimensions. All
 * transforms created from one grid share those buffers, so they must not run
 * concurrently. A grid may be destroyed before its transforms: the buffers are
 * reference counted and released with the last transform. */
#ifndef SPFFT_GRID_FLOAT_H
#define SPFFT_GRID_FLOAT_H

#include "spfft/config.h"
#include "spfft/errors.h"
#include "spfft/types.h"

#ifdef SPFFT_SINGLE_PRECISION
#ifdef __cplusplus
extern "C" {
#endif

typedef void* SpfftFloatGrid;

/* Create a local grid. maxNumLocalZColumns bounds the number of distinct (x,y)
 * columns of any transform created from it. processingUnit: SPFFT_PU_GPU (or
 * SPFFT_PU_HOST | SPFFT_PU_GPU). maxNumThreads < 1 selects the default. */
SPFFT_EXPORT SpfftError spfft_float_grid_create(SpfftFloatGrid* grid, int maxDimX, int maxDimY, int maxDimZ,
                                          int maxNumLocalZColumns,
                                          SpfftProcessingUnitType processingUnit,
                                          int maxNumThreads);

SPFFT_EXPORT SpfftError spfft_float_grid_destroy(SpfftFloatGrid grid);

SPFFT_EXPORT SpfftError spfft_float_grid_max_dim_x(SpfftFloatGrid grid, int* dimX);
SPFFT_EXPORT SpfftError spfft_float_grid_max_dim_y(SpfftFloatGrid grid, int* dimY);
SPFFT_EXPORT SpfftError spfft_float_grid_max_dim_z(SpfftFloatGrid grid, int* dimZ);
SPFFT_EXPORT SpfftError spfft_float_grid_max_num_local_z_columns(SpfftFloatGrid grid,
                                                           int* maxNumLocalZColumns);
SPFFT_EXPORT SpfftError spfft_float_grid_max_local_z_length(SpfftFloatGrid grid, int* maxLocalZLength);
SPFFT_EXPORT SpfftError spfft_float_grid_processing_unit(SpfftFloatGrid grid,
                                                   SpfftProcessingUnitType* processingUnit);
/* Device captured when the grid was created; reselected for every call. */
SPFFT_EXPORT SpfftError spfft_float_grid_device_id(SpfftFloatGrid grid, int* deviceId);
SPFFT_EXPORT SpfftError spfft_float_grid_num_threads(SpfftFloatGrid grid, int* numThreads);

#ifdef __cplusplus
}
#endif
#endif /* SPFFT_SINGLE_PRECISION */

#endif
