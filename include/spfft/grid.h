/* C interface of the grid object (double precision).
 * Signatures follow include/spfft/grid.h:61-175 of the reference; the
 * MPI-only constructor and communicator accessor are not part of this build.
 *
 * A grid owns the device work buffers sized for its maximum dimensions. All
 * transforms created from one grid share those buffers, so they must not run
 * concurrently. A grid may be destroyed before its transforms: the buffers are
 * reference counted and released with the last transform. */
#ifndef SPFFT_GRID_H
#define SPFFT_GRID_H

#include "spfft/config.h"
#include "spfft/errors.h"
#include "spfft/types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void* SpfftGrid;

/* Create a local grid. maxNumLocalZColumns bounds the number of distinct (x,y)
 * columns of any transform created from it. processingUnit: SPFFT_PU_GPU (or
 * SPFFT_PU_HOST | SPFFT_PU_GPU). maxNumThreads < 1 selects the default. */
SPFFT_EXPORT SpfftError spfft_grid_create(SpfftGrid* grid, int maxDimX, int maxDimY, int maxDimZ,
                                          int maxNumLocalZColumns,
                                          SpfftProcessingUnitType processingUnit,
                                          int maxNumThreads);

SPFFT_EXPORT SpfftError spfft_grid_destroy(SpfftGrid grid);

SPFFT_EXPORT SpfftError spfft_grid_max_dim_x(SpfftGrid grid, int* dimX);
SPFFT_EXPORT SpfftError spfft_grid_max_dim_y(SpfftGrid grid, int* dimY);
SPFFT_EXPORT SpfftError spfft_grid_max_dim_z(SpfftGrid grid, int* dimZ);
SPFFT_EXPORT SpfftError spfft_grid_max_num_local_z_columns(SpfftGrid grid,
                                                           int* maxNumLocalZColumns);
SPFFT_EXPORT SpfftError spfft_grid_max_local_z_length(SpfftGrid grid, int* maxLocalZLength);
SPFFT_EXPORT SpfftError spfft_grid_processing_unit(SpfftGrid grid,
                                                   SpfftProcessingUnitType* processingUnit);
/* Device captured when the grid was created; reselected for every call. */
SPFFT_EXPORT SpfftError spfft_grid_device_id(SpfftGrid grid, int* deviceId);
SPFFT_EXPORT SpfftError spfft_grid_num_threads(SpfftGrid grid, int* numThreads);

#ifdef __cplusplus
}
#endif

#endif
