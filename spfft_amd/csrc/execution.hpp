// GPU execution of one transform: device index tables, twiddles, stream and
// the stage sequence (reference: src/execution/execution_gpu.hpp:59-128 and
// src/execution/execution_gpu.cpp:254-410).
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>

#include "grid_internal.hpp"
#include "index_map.hpp"
#include "kernels.hpp"
#include "memory.hpp"
#include "spfft/types.h"

namespace spfft {

template <typename T>
class ExecutionGPU {
public:
  ExecutionGPU(const std::shared_ptr<IndexMap>& map, const std::shared_ptr<GridInternal<T>>& grid);
  ~ExecutionGPU();
  ExecutionGPU(const ExecutionGPU&) = delete;
  ExecutionGPU& operator=(const ExecutionGPU&) = delete;

  // Enqueue on the transform's stream, ordered after the default stream.
  // Pointers may be device or host memory.
  void backward(const T* input, T* output);
  void forward(const T* input, T* output, SpfftScalingType scaling);
  // SYNCHRONOUS: wait for the stream. ASYNCHRONOUS: make the default stream
  // wait for it (reference execution_gpu.cpp:403-410).
  void synchronize(SpfftExecType mode);

  T* space_domain_gpu() const { return grid_->space_gpu(); }
  T* space_domain_host() const { return grid_->space_host(); }
  hipStream_t stream() const { return stream_; }
  // Number of T values in the space domain (2 per element for C2C).
  size_t space_domain_size() const;

private:
  void begin();
  StageArgs<T> base_args() const;

  std::shared_ptr<IndexMap> map_;
  std::shared_ptr<GridInternal<T>> grid_;
  hipStream_t stream_ = nullptr;
  hipEvent_t startEvent_ = nullptr, endEvent_ = nullptr;
  T scale_;
  int numCols_ = 0;
  int zBatch_ = 1, yBatch_ = 1, xBatch_ = 1;
  FftDesc fftX_, fftY_, fftZ_;
  DeviceBuffer<Cx<T>> twX_, twY_, twZ_;
  DeviceBuffer<int> entryValue_, entrySlot_, stickBegin_, stickY_, stickCol_, colX_, colBegin_;
};

}  // namespace spfft
