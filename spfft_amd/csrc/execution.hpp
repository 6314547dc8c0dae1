// GPU execution of one transform: device index tables, twiddles, stream and
// the stage sequence (reference: src/execution/execution_gpu.hpp:59-128 and
// src/execution/execution_gpu.cpp:254-410).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "grid_internal.hpp"
#include "index_map.hpp"
#include "fast_kernels.hpp"
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

  // Enqueue a transform. Pointers may be device or host memory.
  // Single calls run on the default (null) stream itself, which orders them
  // after earlier default-stream work and makes later default-stream work
  // wait for them -- the reference's contract (execution_gpu.cpp:333-344,
  // 403-410) without cross-stream hops. Batched calls (concurrent = true) run
  // on the transform's own stream, fenced against the default stream with
  // events, so independent transforms overlap.
  void backward(const T* input, T* output, bool concurrent = false);
  void forward(const T* input, T* output, SpfftScalingType scaling, bool concurrent = false);
  // SYNCHRONOUS: wait for the work. ASYNCHRONOUS: return; a concurrent call
  // additionally makes the default stream wait for its stream.
  void synchronize(SpfftExecType mode, bool concurrent = false);

  // Batched execution of several transforms of one shape on the fixed-length
  // path: one z launch and one xy launch for the whole batch, on the default
  // stream. The argument tables and the captured launch chain (a HIP graph)
  // are kept for the most recent batches and replayed when the same
  // transforms are called again with the same pointers. Returns false without
  // doing anything when the batch does not qualify (different shapes,
  // generic lengths, host pointers, empty sets).
  static bool run_batch(const std::vector<ExecutionGPU<T>*>& ex, const std::vector<const T*>& in,
                        const std::vector<T*>& out, bool forward,
                        const std::vector<SpfftScalingType>& scaling);

  T* space_domain_gpu() const { return grid_->space_gpu(); }
  T* space_domain_host() const { return grid_->space_host(); }
  hipStream_t stream() const { return stream_; }
  // Number of T values in the space domain (2 per element for C2C).
  size_t space_domain_size() const;

private:
  // Checks for preceding errors and returns the stream to enqueue on.
  hipStream_t begin(bool concurrent);
  StageArgs<T> base_args() const;
  FastArgs<T> fast_args() const;

  std::shared_ptr<IndexMap> map_;
  std::shared_ptr<GridInternal<T>> grid_;
  std::uint64_t serial_ = 0;  // process-unique, never reused
  hipStream_t stream_ = nullptr;
  hipEvent_t startEvent_ = nullptr, endEvent_ = nullptr;
  T scale_;
  int numCols_ = 0;
  int zBatch_ = 1, yBatch_ = 1, xBatch_ = 1;
  FftDesc fftX_, fftY_, fftZ_;
  DeviceBuffer<Cx<T>> twX_, twY_, twZ_;
  DeviceBuffer<int> entryValue_, entrySlot_, stickBegin_, stickY_, stickCol_, colX_, colBegin_;
  // fixed-length C2C path (two kernels per direction)
  bool fast_ = false;
  DeviceBuffer<int> stickAt_;
  DeviceBuffer<int2> zEntries_;
  DeviceBuffer<int4> zInfo_;  // stick runs, when every stick is one run in user order
  DeviceBuffer<int> resSticks_, resBegin_, xCol_, resYM_;
  // whole-plane batch kernels (power-of-two lengths up to 128, C2C): tile pitch
  int pitchT_ = 0;
};

}  // namespace spfft
