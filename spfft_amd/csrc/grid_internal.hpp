// Buffer owner shared by the transforms of one grid
// (reference: src/spfft/grid_internal.hpp:47-141, local constructor
// src/spfft/grid_internal.cpp:49-104).
//
// Device memory of a GPU grid, for maxima (X, Y, Z) and C columns:
//   space : X*Y*Z complex  - space-domain data (real for R2C uses the front half)
//   work  : X*Y*Z complex  - y-stage intermediate [z][y][column] and host staging
//   sticks: Z*C' complex   - z-stage intermediate [z][stick], C' = C rounded up to 16
// Host memory, for grids whose processing unit includes SPFFT_PU_HOST (pinned
// when it also includes SPFFT_PU_GPU, as in reference grid_internal.cpp:85-91):
//   array 1: X*Y*Z complex - host execution's sticks and space domain
//   array 2: X*Y*Z complex - host execution's planes
// Array 1 is also the host mirror of the space domain for GPU transforms; a
// GPU-only grid allocates a pinned mirror on first use instead.
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
  // Host space-domain mirror of X*Y*Z complex values: array 1, or for a
  // GPU-only grid a pinned buffer allocated on first call.
  T* space_host();
  // Pinned staging for host transfers of GPU transforms, at least `bytes`
  // (allocated on first use, grown when a transform needs more).
  void* host_staging(size_t bytes);
  // Host execution buffers (null unless the grid includes SPFFT_PU_HOST).
  T* host_array1() const { return host1_ ? host1_->data() : nullptr; }
  T* host_array2() const { return host2_ ? host2_->data() : nullptr; }

private:
  void allocate();

  int maxDimX_, maxDimY_, maxDimZ_, maxNumLocalZColumns_;
  SpfftProcessingUnitType processingUnit_;
  int numThreads_;
  int deviceId_ = 0;
  DeviceBuffer<T> space_, work_, sticks_;
  std::unique_ptr<PinnedBuffer<T>> host_;
  std::unique_ptr<HostBuffer<T>> host1_, host2_;
  std::unique_ptr<HostBuffer<char>> staging_;
  std::mutex hostMutex_;
};

}  // namespace spfft
