// Buffer owner shared by the transforms of one grid
// (reference: src/spfft/grid_internal.hpp:47-141, local constructor
// src/spfft/grid_internal.cpp:49-104).
//
// Device memory of a GPU grid, for maxima (X, Y, Z) and C columns:
//   space : X*Y*Z complex  - space-domain data (real for R2C uses the front half)
//   work  : X*Y*Z complex  - y-stage intermediate [z][y][column] and host staging
//   sticks: Z*C' complex   - z-stage intermediate [z][stick], C' = C rounded up to 16
// The pinned host mirror of the space domain is allocated on first use.
#pragma once

#include <memory>
#include <mutex>

#include "memory.hpp"
#include "spfft/types.h"

namespace spfft {

template <typename T>
class GridInternal {
public:
  GridInternal(int maxDimX, int maxDimY, int maxDimZ, int maxNumLocalZColumns,
               SpfftProcessingUnitType processingUnit, int numThreads);
  // Deep copy: same parameters, new buffers (reference grid_internal.cpp:233-263).
  GridInternal(const GridInternal& other);
  GridInternal& operator=(const GridInternal&) = delete;

  int max_dim_x() const { return maxDimX_; }
  int max_dim_y() const { return maxDimY_; }
  int max_dim_z() const { return maxDimZ_; }
  int max_num_local_z_columns() const { return maxNumLocalZColumns_; }
  int max_local_z_length() const { return maxDimZ_; }
  SpfftProcessingUnitType processing_unit() const { return processingUnit_; }
  int device_id() const { return deviceId_; }
  int num_threads() const { return numThreads_; }

  T* space_gpu() const { return space_.data(); }
  T* work_gpu() const { return work_.data(); }
  T* sticks_gpu() const { return sticks_.data(); }
  // Pinned host buffer of X*Y*Z complex values, allocated on first call.
  T* space_host();

private:
  void allocate();

  int maxDimX_, maxDimY_, maxDimZ_, maxNumLocalZColumns_;
  SpfftProcessingUnitType processingUnit_;
  int numThreads_;
  int deviceId_ = 0;
  DeviceBuffer<T> space_, work_, sticks_;
  std::unique_ptr<PinnedBuffer<T>> host_;
  std::mutex hostMutex_;
};

}  // namespace spfft
