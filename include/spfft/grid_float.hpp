/* C++ grid class (single precision). Public members follow
 * include/spfft/grid_float.hpp:49-203 of the reference. Copy construction and copy
 * assignment allocate new buffers (deep copy); moves transfer them. */
#ifndef SPFFT_GRID_FLOAT_HPP
#define SPFFT_GRID_FLOAT_HPP

#include <memory>
#include "spfft/config.h"
#include "spfft/transform_float.hpp"
#include "spfft/types.h"

#ifdef SPFFT_SINGLE_PRECISION

namespace spfft {

template <typename T>
class SPFFT_NO_EXPORT GridInternal;

class SPFFT_EXPORT GridFloat {
public:
  GridFloat(int maxDimX, int maxDimY, int maxDimZ, int maxNumLocalZColumns,
       SpfftProcessingUnitType processingUnit, int maxNumThreads);

  GridFloat(const GridFloat&);
  GridFloat(GridFloat&&) = default;
  GridFloat& operator=(const GridFloat&);
  GridFloat& operator=(GridFloat&&) = default;

  /* localZLength must equal dimZ (local grid). */
  TransformFloat create_transform(SpfftProcessingUnitType processingUnit,
                             SpfftTransformType transformType, int dimX, int dimY, int dimZ,
                             int localZLength, int numLocalElements,
                             SpfftIndexFormatType indexFormat, const int* indices) const;

  int max_dim_x() const;
  int max_dim_y() const;
  int max_dim_z() const;
  int max_num_local_z_columns() const;
  int max_local_z_length() const;
  SpfftProcessingUnitType processing_unit() const;
  int device_id() const;
  int num_threads() const;

private:
  std::shared_ptr<GridInternal<float>> grid_;
};

}  // namespace spfft

#endif /* SPFFT_SINGLE_PRECISION */

#endif
